/*
 * ge_physics.h -- C ABI of the MI355X (gfx950) implementation of Galaxy
 * Engine's 2D per-frame physics step.
 *
 * The reference exposes this path as C++ member functions that work in place
 * on std::vector<ParticlePhysics> and a global node vector (citations are
 * path:line under the reference tree, GE/ = GalaxyEngine/).  Each entry point
 * below names the reference function it replaces.  Conventions:
 *   - every call returns an int status: GE_OK (0) or a negative GE_E_* code;
 *     ge_last_error() describes the last failure of a context;
 *   - the context owns all device memory; callers own every host array;
 *   - particle state lives in HBM between calls, stored structure-of-arrays;
 *   - calls are ordered on the context's stream; ge_sync() waits for them;
 *   - one context per host thread (contexts are independent).
 *   - After any sort (ge_sort_by_morton_key, ge_step) particles are stored in
 *     Morton order, as the reference leaves them (morton.cpp:61-62); the id
 *     array travels with them so callers can map results back.
 */
#ifndef GE_PHYSICS_H
#define GE_PHYSICS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GE_OK 0
#define GE_E_INVALID (-1)     /* bad argument */
#define GE_E_NOMEM (-2)       /* device or host allocation failed */
#define GE_E_HIP (-3)         /* HIP runtime error */
#define GE_E_UNSUPPORTED (-4) /* flag this build does not implement */
#define GE_E_CAPACITY (-5)    /* particle count above the context capacity */
#define GE_E_DEVICE (-6)      /* a kernel raised its error flag (see message) */
#define GE_E_NODEVICE (-7)    /* no usable gfx950 device */
#define GE_E_RANGE (-8)       /* coordinates outside what a grid can index */

/* per-particle flag bits; the reference keeps these in ParticleRendering
 * (GE/include/Particles/particle.h:75-140) */
#define GE_FLAG_PINNED 1u
#define GE_FLAG_SPH 2u
#define GE_FLAG_DRAWN 4u
#define GE_FLAG_DARK 8u
#define GE_FLAG_UNIQUE_COLOR 16u /* uniqueColor (read by the colour consumers, ge_render.h) */
#define GE_FLAG_SOLID 32u        /* isSolid (read by the size consumers, ge_render.h) */
#define GE_FLAG_HOT_POINT 64u    /* ParticlePhysics::isHotPoint (ge_thermal.h) */
#define GE_FLAG_SOLIDIFIED 128u  /* ParticlePhysics::hasSolidified (ge_thermal.h) */

typedef struct ge_ctx ge_ctx;

typedef struct {
    uint32_t capacity;    /* maximum particle count (all systems together) */
    int32_t device;       /* HIP device ordinal */
    int32_t enable_sph;   /* allocate SPH state (needed for ge_step with sph=1) */
    uint32_t max_systems; /* independent systems one context can hold (>= 1) */
    int32_t enable_render; /* keep the render fields of ge_render.h on the device */
    int32_t enable_temperature; /* keep the temperature fields of ge_thermal.h (needs enable_sph) */
    uint32_t max_constraints;   /* particle constraints the context can hold (ge_constraints.h) */
    int32_t precise_positions;  /* precision form (no counterpart in the reference): every
                                   position is also kept as a home cell of side 1/16 and a
                                   float offset from that cell's centre; the integrator
                                   drifts the offset, and the walk's particle pairs and node
                                   centres (accuracy >= 1) and the direct sum take separations
                                   from these, so they keep their precision however far the
                                   system sits from the origin (|x|, |y| < 2^20).  pos stays
                                   the nearest float; a position written by anything but the
                                   integrator (upload, constraints, walls) restarts its home
                                   form from that float. */
} ge_config;

/* Structure-of-arrays view of the ParticlePhysics fields on this path
 * (GE/include/Particles/particle.h:5-73).  Upload reads the non-NULL arrays
 * (NULL: zero, or id = index, sph_mass = 1); download writes them. */
typedef struct {
    float *x, *y;         /* pos */
    float *vx, *vy;       /* vel */
    float *ax, *ay;       /* acc */
    float *pvx, *pvy;     /* prevVel */
    float *mass;          /* mass */
    uint32_t *flags;      /* GE_FLAG_* */
    uint32_t *id;         /* id */
    float *sph_mass, *rest_dens, *stiff, *visc, *cohesion, *dens;
    float *press, *press_tmp, *pred_dens;
    float *ppx, *ppy;     /* predPos */
    float *pfx, *pfy;     /* pressF */
} ge_particles;

/* The physics fields of UpdateVariables on this path
 * (GE/include/parameters.h:114-133,311-318). */
typedef struct {
    float G;              /* myVar.G */
    float theta;          /* myVar.theta */
    float softening;      /* myVar.softening */
    float dt_start;       /* timeFactor seen by integrateStart (previous frame) */
    float dt;             /* timeFactor for the rest of the frame */
    int32_t naive;        /* myVar.naive: always use the direct sum */
    int32_t direct_below; /* direct sum when N < this (globalLogic.cpp:1323: 5000) */
    int32_t sph;          /* myVar.isSPHEnabled */
    float sph_h;          /* SPH::radiusMultiplier (SPH.h:17) */
    float sph_cell;       /* NeighborSearchV2AVX2::cellSize (neighborSearch.h:604) */
    float sph_unit_mass;  /* myVar.mass */
    float viscosity;      /* myVar.viscosity */
    float cohesion_coef;  /* myVar.cohesionCoefficient */
    float delta;          /* myVar.delta */
    float stiff_mult;     /* myVar.stiffMultiplier */
    float sph_max_vel;    /* myVar.sphMaxVel */
    int32_t periodic;     /* isPeriodicBoundaryEnabled && !infiniteDomain: positions wrap
                             into the domain (physics.cpp:1104-1110) and every
                             separation takes the nearest image (physics.cpp:117-120,
                             169-172, 417-420) */
    int32_t temperature;  /* isTempEnabled: the coupling of ge_thermal.h (a context
                             created with enable_temperature) */
    int32_t sph_ground;   /* sphGround: SPH::groundModeBoundary after PCISPH
                             (SPH.cpp:283-322) */
    int32_t compensated;  /* Kahan-compensated force accumulation (1 = on) */
    int32_t accuracy;     /* 0: the reference's monopole walk and opening test;
                             1: quadrupole, 2: moments up to the hexadecapole,
                             3: up to the octupole; all about the FP64 centre
                             of mass with the offset-corrected opening test
                             |d| > max(size/theta + |COM - centre|, b_max/theta),
                             and a leaf that any lane of a 64-particle wave
                             opens is summed directly by the whole wave
                             (no counterpart in the reference) */
    uint32_t leaf_capacity; /* 0: the reference's leaf rule (quadtree.cpp:16-17);
                               1..64: a node of at most this many particles is a
                               leaf (autotuned trees; not in the reference) */
    uint32_t walk_block;  /* threads per workgroup of the BH walk: 0 (= 256),
                             64, 128 or 256 */
    int32_t use_graph;    /* ge_step captures its frames in a hipGraph once and
                             replays it while params and layout stay the same */
    int32_t sph_search;   /* SPH candidate sets: 0 NeighborSearchV2AVX2
                             (neighborSearch.h:601-797: floor cells, 3 x 3, no
                             filter); 1 NeighborSearchV2, the reference's path
                             without AVX2 (neighborSearch.h:439-599: truncated
                             cells, ceil(radius / cell) cells each way,
                             |p_j - q| <= sph_search_radius) */
    float sph_search_radius; /* NeighborSearchV2::searchRadius (3.5) */
    float domain_w, domain_h; /* myVar.domainSize (parameters.h:97) */
    float vertical_gravity;   /* myVar.verticalGravity (ground mode) */
    float boundary_friction;  /* myVar.boundaryFriction (ground mode) */
    int32_t constraints;      /* run Physics::constraints after PCISPH (ge_constraints.h) */
    uint32_t reserved[1];
} ge_step_params;

/* Quadtree node as laid out by the reference (GE/include/Physics/quadtree.h:
 * 31-112) in DFS pre-order; child[q] uses q&1 -> +x, q&2 -> +y. */
typedef struct {
    float posx, posy, comx, comy, size, mass;
    uint32_t start, end, next;
    uint32_t child[4];
} ge_node;

/* ---- lifecycle ---------------------------------------------------------- */
void ge_default_config(ge_config *cfg);
/* reference defaults: parameters.h:114-133 and 311-318, SPH.h:17,
 * neighborSearch.h:604, globalLogic.cpp:1323 */
void ge_default_params(ge_step_params *p);
int ge_create(ge_ctx **out, const ge_config *cfg);
int ge_destroy(ge_ctx *ctx);
const char *ge_last_error(const ge_ctx *ctx);
const char *ge_version(void);

/* ---- state transfer ----------------------------------------------------- */
/* replaces the caller filling myParam.pParticles (globalLogic.cpp:3) */
int ge_upload(ge_ctx *ctx, uint32_t n, const ge_particles *p);
/* replaces reading myParam.pParticles back; order is the device order */
int ge_download(ge_ctx *ctx, const ge_particles *p);
int ge_count(const ge_ctx *ctx, uint32_t *n);
/* device memory held by the context (all buffers, fixed at ge_create) */
int ge_memory_bytes(const ge_ctx *ctx, uint64_t *bytes);
/* Batched form (no counterpart in the reference, which steps one scene):
 * nsys independent systems, particles concatenated system by system
 * (counts[s] >= 1 each).  Every stage then runs once for all systems; each
 * system keeps its own bounding box, Morton order and quadtree, and its
 * particles never interact with another system's.  A system's results are
 * bit-identical to stepping it alone.  SPH and the direct-sum branch are
 * single-system only. */
int ge_upload_systems(ge_ctx *ctx, uint32_t nsys, const uint32_t *counts, const ge_particles *p);
/* nsys + 1 offsets of the systems in the device order */
int ge_system_offsets(const ge_ctx *ctx, uint32_t *nsys, uint32_t *offsets, uint32_t cap);
/* Continuous batching (no counterpart in the reference): keep the systems
 * keep[0 .. nkeep) (current indices, ascending), drop the rest, and append
 * nadd new systems of counts[] particles read from p (concatenated; NULL id
 * gives 0 .. count-1 over the appended block).  Kept systems stay resident
 * in HBM and are moved on the device; the new order is the kept systems,
 * then the new ones.  A kept system's later results are the same as if it
 * had been stepped alone. */
int ge_systems_update(ge_ctx *ctx, uint32_t nkeep, const uint32_t *keep, uint32_t nadd,
                      const uint32_t *counts, const ge_particles *p);
/* ge_download restricted to system sys (its particles in device order) */
int ge_download_system(ge_ctx *ctx, uint32_t sys, const ge_particles *p);
int ge_sync(ge_ctx *ctx);
/* non-blocking: *done = 1 when every call queued on the context finished */
int ge_poll(ge_ctx *ctx, int32_t *done);

/* Array-of-structures transfer: Physics::flattenParticles / readFlattenBack
 * (GE/src/Physics/physics.cpp:354-390, 542-565) at the boundary.  The caller's
 * records (e.g. std::vector<ParticlePhysics>, GE/include/Particles/particle.h:
 * 5-73) are read or written in place through byte offsets; off[k] is the
 * offset of the k-th ge_particles field (x, y, vx, ..., pfy: GE_NFIELDS in
 * declaration order) or -1 when the record has no such field.  Flags live in
 * the reference's ParticleRendering, so they come as their own array (NULL:
 * zero on upload, skipped on download).  Download order is the device order,
 * as the reference's vector is after sortParticlesByMortonKey. */
#define GE_NFIELDS 24
typedef struct {
    uint32_t stride;          /* bytes from one record to the next */
    int32_t off[GE_NFIELDS];  /* byte offset of each field, -1 = absent */
} ge_aos_layout;
int ge_upload_aos(ge_ctx *ctx, uint32_t n, const void *records, const ge_aos_layout *layout,
                  const uint32_t *flags);
int ge_download_aos(ge_ctx *ctx, void *records, const ge_aos_layout *layout, uint32_t *flags);

/* ---- one frame ---------------------------------------------------------- */
/* updateScene()'s 2D physics ordering (GE/src/globalLogic.cpp:1155-1372):
 * integrateStart, boundingBox, computeMortonKeys, sortParticlesByMortonKey,
 * Quadtree, acc = 0, calculateForceFromGrid / naiveGravity, pcisphSolver,
 * integrateEnd.  nsteps frames; the first uses dt_start for integrateStart. */
int ge_step(ge_ctx *ctx, const ge_step_params *p, int32_t nsteps);

/* ---- stage entry points (one per reference function) -------------------- */
/* glm::vec3 boundingBox() -- GE/src/globalLogic.cpp:1083-1099 */
int ge_bounding_box(ge_ctx *ctx, float out_bb[3]);
/* Morton::computeMortonKeys -- GE/include/Physics/morton.h:17,
 * GE/src/Physics/morton.cpp:27-38 */
int ge_compute_morton_keys(ge_ctx *ctx, const float bb[3]);
/* Morton::sortParticlesByMortonKey -- GE/include/Physics/morton.h:19-20,
 * GE/src/Physics/morton.cpp:40-63 (stable by (key, index)) */
int ge_sort_by_morton_key(ge_ctx *ctx);
/* Quadtree::Quadtree / Quadtree::root -- GE/include/Physics/quadtree.h:118-128,
 * GE/src/Physics/quadtree.cpp:5-107 */
int ge_build_quadtree(ge_ctx *ctx, const float bb[3]);
/* acc = 0 (globalLogic.cpp:1311-1314), then Physics::calculateForceFromGrid
 * (physics.cpp:91-205) or naiveGravity (physics.cpp:392-434) by the reference's
 * dispatch (globalLogic.cpp:1319-1348) */
int ge_compute_gravity(ge_ctx *ctx, const ge_step_params *p);
/* Physics::integrateStart / integrateEnd -- GE/include/Physics/physics.h:66-68,
 * GE/src/Physics/physics.cpp:1069-1125 */
int ge_integrate_start(ge_ctx *ctx, const ge_step_params *p);
int ge_integrate_end(ge_ctx *ctx, const ge_step_params *p);
/* NeighborSearchV2AVX2::newGridAVX2 -- GE/include/Particles/neighborSearch.h:636 */
int ge_sph_new_grid(ge_ctx *ctx, const ge_step_params *p);
/* SPH::pcisphSolver -- GE/include/Physics/SPH.h:623-632 */
int ge_sph_pcisph(ge_ctx *ctx, const ge_step_params *p);

/* ---- parity hooks ------------------------------------------------------- */
/* Morton keys of the last ge_compute_morton_keys, indexed by particle slot */
int ge_get_keys(ge_ctx *ctx, uint64_t *keys);
/* permutation applied by the last sort: slot i took particle perm[i] */
int ge_get_perm(ge_ctx *ctx, uint32_t *perm);
/* node count and nodes of the last tree (reference layout, pre-order) */
int ge_get_node_count(ge_ctx *ctx, uint32_t *count);
int ge_get_nodes(ge_ctx *ctx, ge_node *nodes, uint32_t cap);
/* BH interaction count (accepted nodes + leaf partners) of the last gravity */
int ge_get_interactions(ge_ctx *ctx, uint64_t *count);
/* the same count split into accepted nodes and leaf partners */
int ge_get_interaction_counts(ge_ctx *ctx, uint64_t *nodes, uint64_t *pairs);
/* walk statistics of the last counted gravity call: {interactions, accepted
 * nodes, node visits summed over waves, leaf partners looped over summed over
 * waves}; lane utilisation = nodes / (64 visits), pairs / (64 leaf trips) */
int ge_get_walk_stats(ge_ctx *ctx, uint64_t out[4]);
/* SPH candidate sets of the current grid (queryNeighborsAVX2,
 * neighborSearch.h:722-777) queried at pos (which=0) or predPos (which=1):
 * CSR offsets off[n+1] and particle slots idx[] (ascending per particle).
 * *total receives the number of entries; idx may be NULL to size it. */
int ge_get_neighbors(ge_ctx *ctx, int32_t which, int64_t *off, uint32_t *idx,
                     int64_t cap, int64_t *total);
/* NeighborSearchV2::newGrid + neighborAmount (neighborSearch.h:481-598):
 * rebuilds the SPH grid over all particles with the NeighborSearchV2
 * semantics (p->sph_cell, p->sph_search_radius) and writes each particle's
 * candidate count (self included) in device order; pinned, drawn and
 * dark-matter particles get 0.  Needs a context with SPH state. */
int ge_neighbor_counts(ge_ctx *ctx, const ge_step_params *p, uint32_t *counts);

/* ---- FP64 diagnostics (no counterpart in the reference, whose UI shows
 * particle counts and FPS only, GE/src/UI/UI.cpp) ----------------------------- */
typedef struct {
    double mass;             /* sum m */
    double px, py;           /* momentum, sum m v */
    double lz;               /* angular momentum about the origin, sum m (x vy - y vx) */
    double kinetic;          /* sum m |v|^2 / 2 */
    double potential;        /* -G sum_{i<j} m_i m_j / sqrt(r^2 + eps^2) (with_potential only) */
    double comx, comy;       /* centre of mass */
} ge_diag;
/* FP64 reductions over system sys of the current state; the potential is an
 * O(N^2) FP64 direct sum with the walk's Plummer softening */
int ge_diagnostics(ge_ctx *ctx, uint32_t sys, const ge_step_params *p, int32_t with_potential,
                   ge_diag *out);
/* ge_diagnostics for callers that cannot allocate memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc); it has
 * no counterpart in the reference.  Same arguments and the same bits as
 * ge_diagnostics, but it calls no hipMalloc, hipFree or allocating library
 * routine: its partial sums live in a fixed device array of 188 KiB per GPU
 * (looked up once per device by ge_create, held until the process ends).
 * Limit: with_potential needs a system of at most 4,194,304 particles
 * (GE_E_CAPACITY beyond it); the moments have no limit.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (contexts used from
 * several host threads at once) race; serialise them. */
int ge_diagnostics_noalloc(ge_ctx *ctx, uint32_t sys, const ge_step_params *p, int32_t with_potential,
                           ge_diag *out);
/* FP64 direct-sum accelerations of the particles in slots targets[k] from
 * every other particle of their system: G m_j d / (|d|^2 + eps^2)^1.5, the
 * force the tree walk approximates (the reference to measure its error) */
int ge_direct_fp64(ge_ctx *ctx, uint32_t ntargets, const uint32_t *targets,
                   const ge_step_params *p, double *ax, double *ay);

/* ---- tracing (no counterpart in the reference, which has only an FPS
 * counter, GE/src/UI/UI.cpp:1317-1324) -------------------------------------- */
/* phases timed with hipEvents on the context stream */
#define GE_PHASE_INTEGRATE_BBOX 0 /* integrateStart + boundingBox */
#define GE_PHASE_KEYS 1           /* computeMortonKeys */
#define GE_PHASE_SORT 2           /* radix sort of the keys */
#define GE_PHASE_GATHER 3         /* permutation gather of the state */
#define GE_PHASE_TREE 4           /* quadtree build */
#define GE_PHASE_GRAVITY 5        /* BH walk or direct sum (+ fused kick) */
#define GE_PHASE_SPH_GRID 6       /* SPH grid builds */
#define GE_PHASE_SPH 7            /* PCISPH passes (+ kick) */
#define GE_PHASE_INTEGRATE_END 8  /* stand-alone integrateEnd */
#define GE_PHASE_MULTIPOLE 9      /* node moments of the accuracy modes */
#define GE_NPHASES 10
/* enable (1) or disable (0) phase timing; clears the accumulators */
int ge_profile_enable(ge_ctx *ctx, int enable);
/* synchronises, then returns accumulated milliseconds and launch counts per
 * phase since the last enable/read, and clears them */
int ge_profile_read(ge_ctx *ctx, double ms[GE_NPHASES], uint32_t launches[GE_NPHASES]);

#ifdef __cplusplus
}
#endif
#endif /* GE_PHYSICS_H */
