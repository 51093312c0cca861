// ge_internal.hpp -- device-side layout and launchers shared by the HIP
// translation units of libge_physics.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ge_physics.h"
#include "../../include/ge_render.h"
#include "../../include/ge_thermal.h"
#include "../../include/ge_constraints.h"
#include "../../include/ge_merger.h"

namespace ge {

constexpr int kBlock = 256;            // threads per workgroup for streaming kernels
constexpr int kIdxBits = 28;           // low bits of a packed sort word hold the slot
constexpr int kMortonBits = 36;        // 18 bits per axis (morton.cpp:3-25)
constexpr uint32_t kChunk = 1024;      // particles per bounding-box reduction chunk (one workgroup each)
constexpr uint64_t kIdxMask = (1ull << kIdxBits) - 1;
constexpr uint32_t kMaxParticles = 1u << kIdxBits;
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Tree build: one launch per level for nodes above kBlockSubMax particles,
// then one workgroup per subtree of at most kBlockSubMax particles, then one
// wave per subtree of at most 64 particles.
constexpr int kTopLevels = 16;         // launches of the grid-wide level phase
constexpr uint32_t kBlockSubMax = 4096;
constexpr int kMaxDepth = 48;          // wave-subtree recursion depth (LDS stack)
constexpr int kBlockLevelsMax = 160;   // local levels a workgroup subtree may take
// Key-run splits of the level phase (one system): for levels L = 1..kKeyLevels
// the first particle (+1, 0 = none) of each run of equal L-digit Morton
// prefixes, and for levels L < kKeyLevels a flag per prefix whose particles'
// float quadrants disagree with their key digits somewhere in it.
constexpr int kKeyLevels = 10;
constexpr uint32_t kExactTag = 1u << 31;  // a node's prefix tag: set bit, then the prefix (at most 2 * kKeyLevels bits)
constexpr uint32_t kKeyStartEntries = ((1u << (2 * (kKeyLevels + 1))) - 4u) / 3u;  // sum of 4^L, L = 1..10
constexpr uint32_t kKeyDirtyEntries = ((1u << (2 * kKeyLevels)) - 1u) / 3u;        // sum of 4^L, L = 0..9
__host__ __device__ constexpr uint32_t key_start_base(int L) { return ((1u << (2 * L)) - 4u) / 3u; }
__host__ __device__ constexpr uint32_t key_dirty_base(int L) { return ((1u << (2 * L)) - 1u) / 3u; }

// device error flag bits
enum : uint32_t {
    kErrNodeCap = 1u,    // final node array too small
    kErrDepth = 2u,      // in-thread recursion deeper than kMaxDepth
    kErrGridRange = 4u,  // SPH grid cannot index the coordinates
    kErrNeighborCap = 8u,
    kErrFrontier = 16u,  // workgroup subtree frontier or level bound exceeded
    kErrMultipole = 32u, // multipole work lists overflowed
    kErrSort = 64u,      // a radix-sort look-back wait passed its spin limit
    kErrConsCap = 128u,  // created constraints beyond max_constraints (the excess is dropped)
};

// Text for a set of device error flags ("node capacity tree depth ..."), used
// by the 2D and 3D contexts' reports.
inline void describe_device_errors(uint32_t e, char *buf, size_t len) {
    std::snprintf(buf, len, "device error flags 0x%x (%s%s%s%s%s%s%s%s)", e, (e & kErrNodeCap) ? "node capacity " : "",
                  (e & kErrDepth) ? "tree depth " : "", (e & kErrGridRange) ? "SPH grid range " : "",
                  (e & kErrNeighborCap) ? "neighbour cap " : "", (e & kErrFrontier) ? "tree frontier " : "",
                  (e & kErrMultipole) ? "multipole lists " : "", (e & kErrSort) ? "sort wait " : "",
                  (e & kErrConsCap) ? "constraint capacity" : "");
}

// counters block (device, uint32 each)
enum CounterSlot : int {
    kCntError = 0,
    kCntBboxTicket = 1,
    kCntPredTicket = 2,
    kCntNodeTotal = 3,      // final node count (root subtree size)
    kCntTopTotal = 4,       // number of top nodes
    kCntTopAlloc = 5,       // next free top node (workgroup-subtree phase)
    kCntBlockRoots = 6,     // number of workgroup-subtree roots
    kCntStamp = 7,          // tree builds so far (stamps nodes finished by the block mass pass)
    kCntLevel0 = 8,         // kCntLevel0 + L: nodes allocated on level L
    kCntLevelBase0 = 8 + 64,  // first top index of level L
    kCntBlistAlloc = 8 + 128, // entries of the per-subtree internal-node lists
    kCntBlistOverflow = 8 + 129, // a workgroup subtree's list did not fit
    kCntWalkQ0 = 8 + 136,   // kWalkQueues tickets of the persistent walk, then its finished-wave count
    kCntSlots = 8 + 146,
};

// ticket ranges of a persistent walk launch (one per XCD)
constexpr uint32_t kWalkQueues = 8;

// Periodic domain of the reference's isPeriodicBoundaryEnabled && !infiniteDomain
// (w = 0: no wrap).
struct Domain {
    float w = 0.0f, h = 0.0f;
};

// d -= size * ((d > size/2) - (d < -size/2))  (physics.cpp:117-120, 169-172, 417-420)
__device__ __forceinline__ float wrap_d(float d, float size, float half) {
    return d - size * (float)((d > half) - (d < -half));
}

// One buffer set of particle state (structure of arrays).
struct State {
    float *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr;
    float *ax = nullptr, *ay = nullptr, *pvx = nullptr, *pvy = nullptr, *m = nullptr;
    uint32_t *flags = nullptr, *id = nullptr;
    float *sph_mass = nullptr, *rest_dens = nullptr, *stiff = nullptr, *visc = nullptr;
    float *cohesion = nullptr, *dens = nullptr, *press = nullptr, *press_tmp = nullptr;
    float *pred_dens = nullptr, *ppx = nullptr, *ppy = nullptr, *pfx = nullptr, *pfy = nullptr;
    // render fields (allocated with enable_render): RGBA8 colours, size, neighbour count
    uint32_t *color = nullptr, *pcolor = nullptr, *scolor = nullptr;
    float *size = nullptr, *prev_size = nullptr;
    int32_t *nbrs = nullptr;
    // temperature fields (allocated with enable_temperature)
    float *temp = nullptr, *ke = nullptr, *prev_ke = nullptr;
    uint32_t *label = nullptr;
    // precision form (allocated with precise_positions): home cell and offset
    int32_t *hx = nullptr, *hy = nullptr;
    float *ox = nullptr, *oy = nullptr;
};

// ---- precision form: a position is kept as a home cell h (side kHomeCell,
// centre h * kHomeCell) and a float offset o from that centre, |o| <= cell/2
// after each drift.  Its float view, the nearest float to h * cell + o, is
// what every other part of the frame reads as pos.  Separations taken as
// (h_j - h_i) * cell + (o_j - o_i) are exact up to one rounding of the
// result, wherever the two positions sit (|pos| < 2^20 keeps h * cell exact).
constexpr float kHomeCell = 0.0625f;
constexpr float kHomeInv = 16.0f;

__device__ __forceinline__ float home_view(int32_t h, float o) {
#pragma clang fp contract(off)
    return (float)h * kHomeCell + o;  // the product is exact
}

// A position whose view is not x was written by something other than the
// integrator (upload, constraints, walls): its home form restarts from x.
__device__ __forceinline__ void home_sync(float x, int32_t &h, float &o) {
#pragma clang fp contract(off)
    if (home_view(h, o) == x) return;
    const float c = rintf(x * kHomeInv);
    if (!(fabsf(c) < 2147483520.0f)) {  // NaN or beyond int32 cells: x is the offset
        h = 0;
        o = x;
        return;
    }
    h = (int32_t)c;
    o = x - c * kHomeCell;  // exact: within half a cell of x
}

// whole cells moved from the offset into the home cell
__device__ __forceinline__ void home_norm(int32_t &h, float &o) {
#pragma clang fp contract(off)
    const float c = rintf(o * kHomeInv);
    if (c != 0.0f && fabsf(c) < 1073741824.0f) {
        h += (int32_t)c;
        o = o - c * kHomeCell;  // exact
    }
}

// the position moved by a float distance d (the periodic wrap): whole cells
// of d go to the home cell, the rest (exact, at most half a cell) to the offset
__device__ __forceinline__ void home_shift(int32_t &h, float &o, float d) {
#pragma clang fp contract(off)
    const float c = rintf(d * kHomeInv);
    h += (int32_t)c;
    o = o + (d - c * kHomeCell);
    home_norm(h, o);
}

// separation b - a of two home positions
__device__ __forceinline__ float home_sep(int32_t hb, float ob, int32_t ha, float oa) {
#pragma clang fp contract(off)
    return (float)(hb - ha) * kHomeCell + (ob - oa);
}

// Node record read by the traversal: 32 bytes, one scalar load per visit.
struct alignas(16) NodeRec {
    float comx, comy, mass, size_sq;
    uint32_t next, start, end, child_mask;
};

// Top part of the tree, built level by level (structure of arrays).
struct TopNodes {
    float *px = nullptr, *py = nullptr, *size = nullptr;
    uint32_t *start = nullptr, *end = nullptr;
    uint32_t *slot = nullptr;      // scratch per node: its entry in the top-mass pass's list
    uint32_t *exact = nullptr;     // level phase: kExactTag | P if the node's range is the run of key prefix P, else 0
    uint32_t *child = nullptr;     // 4 per node
    uint32_t *kind = nullptr;      // 0 internal, 1 leaf, 2 wave subtree, 3 pending, 4 dead
    uint32_t *pending2 = nullptr;  // build stamp of nodes whose mass the block pass computed
    uint32_t *offset = nullptr;    // pre-order index in the final array
    uint32_t *broot = nullptr;     // roots handed to the workgroup-subtree phase
    uint32_t *blist = nullptr;     // internal nodes of each workgroup subtree, level by level
    uint32_t *blist_off = nullptr; // per workgroup-subtree root: first entry in blist
    uint32_t *blist_cnt = nullptr; // and its entry count (kNone: not listed)
    uint32_t *depth = nullptr;     // tree level of the node
    uint32_t cap = 0;
};

// SPH uniform grid: particles sorted by cell over the occupied cell range;
// first[c] = number of particles in cells before c (W*H + 1 entries), so the
// particles of cells c0..c1 of one row are positions [first[c0], first[c1+1]).
struct Grid {
    uint64_t *keys = nullptr;      // packed (cell << 28 | slot), sorted
    uint64_t *keys_alt = nullptr;
    uint32_t *first = nullptr;     // prefix cell table
    uint32_t *ord = nullptr;       // slot of the k-th particle in cell order
    int32_t *frame = nullptr;      // device: {cx0, cy0, W, H, valid} of the table
    float4 *pk0 = nullptr;         // per cell-order position: packed fields
    float4 *pk1 = nullptr;
    uint32_t *pkf = nullptr;       // flags in cell order
    float2 *xy = nullptr;          // the position the grid was built on, NaN for particles no pass visits
    uint32_t cells_cap = 0;
};
constexpr int kCellBits = 24;      // cell keys below 2^24

struct Buffers {
    State st[2];
    int cur = 0;
    float4 *pxym = nullptr;        // sorted (x, y, m, 0) for the traversal
    float4 *pho = nullptr;         // precision form: sorted (ox, oy, hx, hy) (cells as int bits)
    uint64_t *sort_a = nullptr, *sort_b = nullptr;  // packed sort words
    uint64_t *keys = nullptr;      // 36-bit Morton key per slot (pre-sort order)
    uint32_t *perm = nullptr;      // last permutation
    uint32_t *hist = nullptr;      // radix histogram table
    uint32_t hist_cap = 0;
    float *partials = nullptr;     // bbox partial min/max, 4 per block
    float *partials2 = nullptr;    // second reduction (predPos)
    uint32_t *counters = nullptr;  // CounterSlot
    float *bb = nullptr;           // device bbox {x, y, size, _} and predPos box at +4
    uint64_t *interactions = nullptr;
    // systems (a single scene is one system)
    uint32_t *sys_off = nullptr;   // nsys + 1 particle offsets
    uint32_t *sys_of = nullptr;    // system of each particle slot
    float *sys_bb = nullptr;       // 8 per system: {x, y, size, 0, minx, miny, maxx, maxy}
    uint32_t *chunk_lo = nullptr;  // reduction chunks: first slot of each chunk (+ end)
    uint32_t *chunk_first = nullptr;  // first chunk of each system (nsys + 1)
    uint32_t *sys_src = nullptr;      // compaction: old first slot of each kept system
    uint32_t *sys_wave_off = nullptr; // nsys + 1 prefix of walk waves per system (64 particles each)
    uint32_t *ccreate = nullptr;      // constraint creation: per-slot counts, block sums, totals
    uint32_t max_chunks = 0;
    TopNodes top;
    uint32_t *start_count = nullptr;  // per particle: nodes whose range starts there (then its prefix)
    uint32_t *kstart = nullptr;       // key-run starts of the top levels (kKeyStartEntries)
    uint8_t *kdirty = nullptr;        // prefixes whose float quadrants disagree (kKeyDirtyEntries)
    uint32_t *start_level = nullptr;  // per particle: shallowest level of such a node
    uint32_t *scan_part = nullptr;    // block sums of the node-start scan
    NodeRec *nodes = nullptr;
    float4 *node_geom = nullptr;   // (posx, posy, size, 0) per final node
    float4 *node_mp = nullptr;     // accuracy modes: 6 float4 multipole record per node
    float4 *node_home = nullptr;   // precision form: expansion centre per node as (ox, oy, hx, hy)
    uint32_t *mp_cnt = nullptr;    // accuracy modes: list counters of the moment pass
    uint32_t *mp_wave_list = nullptr, *mp_big_list = nullptr, *mp_piece_owner = nullptr;
    double *mp_piece_sum = nullptr;
    uint32_t mp_wave_cap = 0, mp_big_cap = 0, mp_piece_cap = 0;
    uint32_t node_cap = 0;
    float *sphf = nullptr;         // SPH force scratch (x, y interleaved), 2 per particle
    float *sph_pp = nullptr;       // grid-A order (predX, predY, restDens, sphMass)
    float *sph_pd = nullptr;       // grid-A order (press, predDens)
    float *sph_ppxy = nullptr;     // grid-A order (predX, predY) of visitors, NaN otherwise
    Grid grid_pos, grid_pred;
    // temperature coupling (enable_temperature): Node::gridTemp per final
    // node, the bottom-up pass's parent links and arrival counts, the
    // material table
    float *node_temp = nullptr;
    float *node_tavg = nullptr;  // gridTemp / particle count of each node (what the walk reads)
    uint32_t *node_parent = nullptr, *node_arrive = nullptr;
    ge_material *materials = nullptr;
    // particle constraints (max_constraints): the constraints and their
    // end indices (double-buffered for the compaction), colours, keep flags,
    // block counts, device count / frame counter / colouring counter, the
    // colour order and colour starts, sort words and their histogram
    ge_constraint *cons = nullptr, *cons_tmp = nullptr;
    uint2 *cidx = nullptr, *cidx_tmp = nullptr;
    uint32_t *ccolor = nullptr, *ccolor_tmp = nullptr, *ckeep = nullptr, *cblock = nullptr;
    uint32_t *ccount = nullptr, *cframe = nullptr, *cleft = nullptr;
    uint32_t *corder = nullptr, *cstart = nullptr;
    uint64_t *cw_a = nullptr, *cw_b = nullptr;
    uint32_t *chist = nullptr;
};

// merger work buffers, allocated on the first ge_merge
struct MergerBuf {
    uint64_t *wa = nullptr, *wb = nullptr;
    uint32_t *hist = nullptr;
    float *rad = nullptr;
    uint32_t *del = nullptr, *ecnt = nullptr, *eoff = nullptr, *edges = nullptr, *parent = nullptr;
    float4 *part = nullptr;
    uint32_t *grid = nullptr, *flag = nullptr, *dlist = nullptr, *bc = nullptr, *perm = nullptr;
    size_t edge_cap = 0;
    uint32_t cap = 0;
};

}  // namespace ge

struct ge_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cap = 0;
    uint32_t n = 0;
    uint32_t max_sys = 1;
    uint32_t nsys = 1;
    uint32_t nchunks = 1;
    uint32_t nwaves = 0;  // walk waves: each system's particles in 64-particle waves from its first slot
    uint32_t cus = 256;   // compute units of the device (sizes persistent launches)
    int idx_bits = 28;             // low bits of a packed sort word: slot within the system
    int sys_bits = 0;              // high bits: system index
    std::vector<uint32_t> host_off;  // system offsets (host copy)
    bool sph = false;
    bool render = false;           // render fields allocated
    bool thermal = false;          // temperature fields allocated
    bool precise = false;          // precision form (home cells) allocated
    ge_thermal_params thermal_p{};  // ge_set_thermal
    uint32_t nmat = 0;             // materials in b.materials
    uint32_t cons_cap = 0;         // max_constraints
    uint32_t cons_bound = 0;       // upper bound of the device constraint count
    bool cons_colored = true;      // colours match the constraints
    uint32_t ncolors = 0;
    ge_constraint_params cons_p{};
    bool debug = true;             // keep node geometry for ge_get_nodes
    bool have_tree = false;
    bool have_grid = false;
    float sph_cell = 3.0f;         // cell size grid A was built with
    int sph_search = 0;            // search semantics grid A was built with
    float sph_radius = 3.5f;       // and its search radius
    uint32_t leaf_cap = 0;         // leaf rule of the next tree build (0 = reference)
    uint64_t device_bytes = 0;     // device memory the context allocated
    double *diag = nullptr;        // FP64 diagnostics scratch (allocated on first use)
    size_t diag_cap = 0;
    uint64_t layout_version = 0;   // bumped whenever the particle layout changes
    // hipGraph of two frames (the state and sort buffers alternate, so a pair
    // returns every buffer pointer to where it started)
    struct FrameGraph {
        hipGraphExec_t exec = nullptr;
        ge_step_params params{};
        uint64_t layout_version = 0;
        int cur = -1;
        const void *sort_a = nullptr, *grid_a = nullptr, *grid_b = nullptr;
    } graph;
    ge::Buffers b;
    ge::MergerBuf merger;
    std::string err;
    // phase timing (ge_profile_enable)
    bool prof = false;
    struct PhaseEvent { int phase; hipEvent_t a, b; };
    std::vector<PhaseEvent> prof_events;
};

namespace ge {

// ---- launchers (each defined next to its kernels) ----
// sort.hip
// err: the context's error word (counters + kCntError), where a look-back
// wait that passed its spin limit raises kErrSort; every caller passes it so
// that the next ge_sync reports a sort that gave up
void radix_sort_words(hipStream_t stream, uint32_t *hist, uint64_t *&data, uint64_t *&alt, uint32_t n, int lo_bit,
                      int hi_bit, uint32_t *err);
void radix_sort_packed(ge_ctx *c, uint64_t *&data, uint64_t *&alt, uint32_t n,
                       int lo_bit, int hi_bit);
void radix_sort_kv(hipStream_t stream, uint32_t *hist, uint64_t *&keys, uint64_t *&keys_alt, uint32_t *&vals,
                   uint32_t *&vals_alt, uint32_t n, int lo_bit, int hi_bit, uint32_t *err);
uint32_t radix_hist_entries(uint32_t cap);
// diag.hip: looks up the device address of the fixed scratch of
// ge_diagnostics_noalloc for `device` (current), once per device
int noalloc_setup(int device);
// step.hip
void launch_integrate_start_bbox(ge_ctx *c, float dt, int sph_clamp, float max_vel, int do_integrate,
                                 Domain dom = Domain{});
Domain domain_of(const ge_step_params *p);
void launch_sph_ground(ge_ctx *c, const ge_step_params *p);
void launch_morton_keys(ge_ctx *c);
void launch_sort_keys(ge_ctx *c);
void launch_gather(ge_ctx *c);
void launch_bh(ge_ctx *c, const ge_step_params *p, float end_kick_dt, bool count);
int ensure_multipole_buffers(ge_ctx *c);
void launch_multipoles(ge_ctx *c, float theta);
void launch_direct(ge_ctx *c, const ge_step_params *p, float end_kick_dt);
void launch_zero_acc(ge_ctx *c);
void launch_integrate_end(ge_ctx *c, float dt);
void launch_sys_of(ge_ctx *c);
void launch_compact(ge_ctx *c, uint32_t n_kept);
// tree.hip
// sorted_words: the packed sort words in b.sort_a are the sort of the
// current particle order (the frame's own sort), so the level phase may split
// Morton-prefix runs without searching
void launch_tree_build(ge_ctx *c, bool sorted_words = false);
// sph.hip
int sph_new_grid(ge_ctx *c, const ge_step_params *p);
int sph_pcisph(ge_ctx *c, const ge_step_params *p, float end_kick_dt);
int sph_neighbors(ge_ctx *c, int which, int64_t *off, uint32_t *idx, int64_t cap, int64_t *total);
int sph_neighbor_counts(ge_ctx *c, const ge_step_params *p, uint32_t *counts);
int sph_count_into(ge_ctx *c, int search, float cell, float radius, int32_t *out);
int sph_create_constraints(ge_ctx *c, const ge_step_params *p, uint32_t *created);
// thermal.hip
void launch_thermal_fill(ge_ctx *c, const State &s, uint32_t at, uint32_t cnt);
void launch_node_temps(ge_ctx *c);
void launch_temperature(ge_ctx *c, const ge_step_params *p);
// merger.hip
void merger_free(ge_ctx *c);
int merger_cell_size(ge_ctx *c, const ge_merger_params *mp, float *cell);
int merger_run(ge_ctx *c, const ge_merger_params *mp, std::vector<uint32_t> &dels);
int merger_apply(ge_ctx *c, const std::vector<uint32_t> &dels, uint32_t &new_n);
// constraints.hip
int cons_prepare(ge_ctx *c);
int cons_solve(ge_ctx *c, const ge_step_params *p, bool map_done);
// render.hip
void launch_render_fill(hipStream_t st, uint32_t *color, uint32_t *pcolor, uint32_t *scolor, float *size,
                        float *prev_size, int32_t *nbrs, uint32_t at, uint32_t cnt, uint32_t rgba, float sz);
void launch_spawn_galaxy(ge_ctx *c, const State &s, uint32_t at, const ge_galaxy_params *g, uint64_t seed,
                         const float *d_uniforms);

inline uint32_t blocks_for(uint64_t n, int per_block) {
    uint64_t b = (n + per_block - 1) / per_block;
    return b == 0 ? 1u : (uint32_t)b;
}

// glm::min / glm::max scalar forms used by the reference bbox
__device__ __forceinline__ float glm_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float glm_max(float a, float b) { return (a < b) ? b : a; }

// XCD-aware block order: consecutive logical blocks (neighbouring Morton
// ranges) are placed on the same XCD so its L2 sees one spatial region.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    uint32_t per = nblocks / kXcd;
    uint32_t full = per * kXcd;
    if (bid >= full) return bid;  // tail blocks keep their id
    return (bid % kXcd) * per + bid / kXcd;
}

__device__ __forceinline__ uint32_t atomic_load_u32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float atomic_load_f32(const float *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace ge

#define GE_HIP_CHECK(c, expr)                                                  \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            (c)->err = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return GE_E_HIP;                                                   \
        }                                                                      \
    } while (0)
