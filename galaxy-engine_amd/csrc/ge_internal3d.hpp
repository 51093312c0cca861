// ge_internal3d.hpp -- device-side layout of the 3D context (ge3_ctx)
// shared by ge_step3d.hip, ge_octree.hip, ge_sph3d.hip and ge_abi3d.hip.
// Not part of the public ABI.
#pragma once

#include "ge_internal.hpp"
#include "../../include/ge_physics3d.h"

namespace ge {

constexpr int kMortonBits3 = 63;  // 21 bits per axis (morton.cpp:67-92)

// Periodic cube of the reference's 3D wrap, centred on the origin (w = 0: none).
struct Domain3 {
    float w = 0.0f, h = 0.0f, d = 0.0f;
};

struct State3 {
    float *x = nullptr, *y = nullptr, *z = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr;
    float *ax = nullptr, *ay = nullptr, *az = nullptr;
    float *pvx = nullptr, *pvy = nullptr, *pvz = nullptr;
    float *m = nullptr;
    uint32_t *flags = nullptr, *id = nullptr;
    // SPH3D fields (allocated with enable_sph)
    float *sph_mass = nullptr, *rest_dens = nullptr, *stiff = nullptr, *visc = nullptr;
    float *cohesion = nullptr, *dens = nullptr, *press = nullptr, *press_tmp = nullptr;
    float *pred_dens = nullptr, *ppx = nullptr, *ppy = nullptr, *ppz = nullptr;
    float *pfx = nullptr, *pfy = nullptr, *pfz = nullptr;
    // render fields (allocated with enable_render): RGBA8 colours, size, neighbour count
    uint32_t *color = nullptr, *pcolor = nullptr, *scolor = nullptr;
    float *size = nullptr, *prev_size = nullptr;
    int32_t *nbrs = nullptr;
};

// SPH3D neighbour grid: particles sorted by a hash slot of their cell,
// first[s] the first sorted position of slot s (s in [0, slots]); every
// query compares exact cell coordinates, so the candidate sets are the
// reference's whatever the hash.
struct Grid3 {
    uint64_t *keys = nullptr, *keys_alt = nullptr;  // (slot << kIdxBits | particle), sorted
    uint32_t *first = nullptr;
    uint32_t *ord = nullptr;   // particle of sorted position k
    int4 *cell = nullptr;      // (cx, cy, cz, flags) per sorted position
    float4 *pk0 = nullptr;     // grid A: (x, y, z, sphMass); grid B: (ppx, ppy, ppz, press)
    float4 *pk1 = nullptr;     // grid A: (vx, vy, vz, visc); grid B: (predDens, sphMass, -, -)
    float2 *pk2 = nullptr;     // grid A: (cohesion, dens)
};

// Octree node record read by the walk: 32 bytes, two 16-byte scalar loads.
struct alignas(16) NodeRec3 {
    float comx, comy, comz, mass;
    float size_sq;
    uint32_t next, start, end;
};

// Top part of the octree, built level by level (structure of arrays).
struct TopNodes3 {
    float *px = nullptr, *py = nullptr, *pz = nullptr, *size = nullptr;
    uint32_t *start = nullptr, *end = nullptr;
    uint32_t *child = nullptr;     // 8 per node
    uint32_t *kind = nullptr;      // the kinds of the quadtree build
    uint32_t *pending2 = nullptr;  // build stamp of nodes whose mass the block pass computed
    uint32_t *offset = nullptr;    // pre-order index in the final array
    uint32_t *broot = nullptr;     // roots handed to the workgroup-subtree phase
    uint32_t *blist = nullptr;     // internal nodes of each workgroup subtree, level by level
    uint32_t *blist_off = nullptr;
    uint32_t *blist_cnt = nullptr;
    uint32_t *depth = nullptr;
    uint32_t *exact = nullptr;     // level phase: kExactTag | P if the node's range is the run of key prefix P, else 0
    uint32_t cap = 0;
};

// Key-run splits of the octree level phase: for levels L = 1..kKeyLevels3
// the first particle (+1, 0 = none) of each run of equal L-digit (3-bit
// digit) Morton prefixes, and for levels L < kKeyLevels3 a flag per prefix
// whose particles' float octant tests disagree with their key digits.
constexpr int kKeyLevels3 = 7;
constexpr uint32_t kKeyStartEntries3 = ((1u << (3 * (kKeyLevels3 + 1))) - 8u) / 7u;  // sum of 8^L, L = 1..7
constexpr uint32_t kKeyDirtyEntries3 = ((1u << (3 * kKeyLevels3)) - 1u) / 7u;        // sum of 8^L, L = 0..6
__host__ __device__ constexpr uint32_t key_start_base3(int L) { return ((1u << (3 * L)) - 8u) / 7u; }
__host__ __device__ constexpr uint32_t key_dirty_base3(int L) { return ((1u << (3 * L)) - 1u) / 7u; }

struct Buffers3 {
    State3 st[2];
    int cur = 0;
    float4 *pxyzm = nullptr;       // sorted (x, y, z, m) for the walk
    uint64_t *keys = nullptr;      // 63-bit key per slot (pre-sort order)
    uint64_t *sk_a = nullptr, *sk_b = nullptr;  // sort keys
    uint32_t *sv_a = nullptr, *sv_b = nullptr;  // sort values (slots)
    uint32_t *perm = nullptr;      // last permutation
    uint32_t *hist = nullptr;
    float *partials = nullptr;     // bbox partial min/max, 8 per block
    float *bb = nullptr;           // device {x, y, z, size, minx, miny, minz, 0, maxx, maxy, maxz, 0}
    uint32_t *counters = nullptr;  // CounterSlot of ge_internal.hpp
    unsigned long long *inter = nullptr;  // interactions, accepted nodes
    TopNodes3 top;
    uint32_t *start_count = nullptr;
    uint32_t *kstart = nullptr;       // key-run starts of the top levels (kKeyStartEntries3)
    uint8_t *kdirty = nullptr;        // prefixes whose float octant tests disagree (kKeyDirtyEntries3)
    uint32_t *start_level = nullptr;
    uint32_t *scan_part = nullptr;
    NodeRec3 *nodes = nullptr;
    float4 *node_geom = nullptr;   // (posx, posy, posz, size) per final node
    uint32_t *node_mask = nullptr; // occupied octants per final node
    uint32_t node_cap = 0;
    double *diag = nullptr;        // FP64 direct-sum scratch (on first use)
    size_t diag_cap = 0;
    // SPH3D (enable_sph)
    Grid3 grid_pos, grid_pred;     // over pos (A) and over predPos of visiting particles (B)
    float4 *sphf = nullptr;        // pair-force sums per particle
    float4 *sph_pp = nullptr;      // grid-A order: (ppx, ppy, ppz, restDens)
    float2 *sph_pd = nullptr;      // grid-A order: (press, predDens)
    uint32_t slot_bits = 0;        // hash slots = 1 << slot_bits
    // neighbour lists of the SPH passes (density and own pressure visits at
    // predPos, reverse pressure visits): nbr_cap entries per particle,
    // column-major, and a count per particle
    uint32_t *nbr_idx[2] = {nullptr, nullptr};
    uint32_t *nbr_cnt[2] = {nullptr, nullptr};
    uint32_t nbr_cap = 0;
};

}  // namespace ge

struct ge3_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cap = 0;
    uint32_t n = 0;
    uint32_t nblocks_bb = 1;       // workgroups of the bbox reduction
    uint32_t leaf_cap = 0;
    bool have_tree = false;
    bool sph = false;              // SPH3D state allocated
    bool render = false;           // render fields allocated
    bool have_grid = false;        // grid A matches the current positions
    float grid_cell = 3.0f;        // cell size grid A was built with
    ge::Buffers3 b;
    std::string err;
};

namespace ge {

// ge_step3d.hip
void launch3_integrate_bbox(ge3_ctx *c, float dt, int do_integrate, Domain3 dom, int sph_clamp = 0,
                           float sph_max_vel = 0.0f);
void launch3_morton_keys(ge3_ctx *c, const float *bb_device);
void launch3_sort_gather(ge3_ctx *c);
void launch3_zero_acc(ge3_ctx *c);
void launch3_bh(ge3_ctx *c, const ge3_step_params *p, float end_kick_dt, bool count);
void launch3_direct(ge3_ctx *c, const ge3_step_params *p, float end_kick_dt);
void launch3_integrate_end(ge3_ctx *c, float dt);
void launch3_direct_fp64(ge3_ctx *c, double G, double eps, uint32_t nt, const uint32_t *targets, double *out);
void launch3_pack(ge3_ctx *c);
// ge_octree.hip
// sorted_keys: the sort keys in b.sk_a are the sort of the current particle
// order (the frame's own sort), so the level phase may split Morton-prefix
// runs without searching
void launch_octree_build(ge3_ctx *c, const float *bb_device, bool sorted_keys = false);
// ge_sph3d.hip
void sph3_new_grid(ge3_ctx *c, const ge3_step_params *p);
void sph3_pcisph(ge3_ctx *c, const ge3_step_params *p, float end_kick_dt);
void sph3_ground(ge3_ctx *c, const ge3_step_params *p);
int sph3_neighbors(ge3_ctx *c, int which, int64_t *off, uint32_t *idx, int64_t cap, int64_t *total);
int sph3_neighbor_counts(ge3_ctx *c, uint32_t *counts);
void sph3_count_into(ge3_ctx *c, float cell, int32_t *out);

constexpr uint32_t kBBoxChunk3 = 1024;  // particles per bbox workgroup

}  // namespace ge
