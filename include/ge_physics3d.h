/*
 * ge_physics3d.h -- C ABI of the MI355X (gfx950) implementation of Galaxy
 * Engine's 3D physics step: mode3D() (GE/src/globalLogic.cpp:1552-1722)
 * with Morton3D keys, the Octree, Physics3D and SPH3D.  Same conventions as
 * ge_physics.h: int status codes (GE_OK or a
 * negative GE_E_*), a context that owns all device memory, state resident in
 * HBM as structure-of-arrays, calls ordered on the context's stream, one
 * context per host thread, Morton order after every sort with `id`
 * travelling along.  GE/ = GalaxyEngine/.
 *
 * Not implemented (GE_E_UNSUPPORTED, the caller keeps its CPU branch):
 * temperature transfer (isTempEnabled).  SPH3D follows the AVX2 neighbour
 * search (NeighborSearch3DV2AVX2, GE/include/Particles/neighborSearch.h:
 * 978-1193), the one an x86-64 build with -mavx2 takes.
 */
#ifndef GE_PHYSICS3D_H
#define GE_PHYSICS3D_H

#include <stdint.h>

#include "ge_physics.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ge3_ctx ge3_ctx;

typedef struct {
    uint32_t capacity;  /* maximum particle count */
    int32_t device;     /* HIP device ordinal */
    int32_t enable_sph; /* allocate SPH3D state (needed for sph = 1) */
    int32_t enable_render; /* keep the render fields of ge_render.h on the device */
    uint32_t reserved[4];
} ge3_config;

/* Structure-of-arrays view of ParticlePhysics3D on this path
 * (GE/include/Particles/particle.h:142-212).  Upload reads the non-NULL
 * arrays (NULL: zero, id = index, sph_mass = 1); download writes them.  The
 * SPH fields are read and written only by a context created with
 * enable_sph. */
typedef struct {
    float *x, *y, *z;      /* pos */
    float *vx, *vy, *vz;   /* vel */
    float *ax, *ay, *az;   /* acc */
    float *pvx, *pvy, *pvz; /* prevVel */
    float *mass;
    uint32_t *flags;       /* GE_FLAG_* (pinned, SPH, being drawn, dark matter) */
    uint32_t *id;
    float *sph_mass, *rest_dens, *stiff, *visc, *cohesion, *dens;
    float *press, *press_tmp, *pred_dens;
    float *ppx, *ppy, *ppz; /* predPos */
    float *pfx, *pfy, *pfz; /* pressF */
} ge3_particles;

/* The UpdateVariables fields the 3D gravity frame reads
 * (GE/include/parameters.h:99,114-119,133,138,318). */
typedef struct {
    float G;               /* myVar.G */
    float theta;           /* myVar.theta */
    float softening;       /* myVar.softening, added unsquared to r^2 in 3D
                              (physics3D.cpp:166, 213, 451) */
    float dt_start;        /* timeFactor seen by integrateStart3D; mode3D uses
                              the frame's own timeFactor, so dt_start = dt */
    float dt;              /* myVar.timeFactor */
    int32_t naive;         /* myVar.naive: always the direct sum */
    int32_t direct_below;  /* direct sum when N < this (globalLogic.cpp:1678: 5000) */
    int32_t periodic;      /* isPeriodicBoundaryEnabled && !infiniteDomain: positions
                              wrap into [-d/2, d/2) per axis (physics3D.cpp:782-800)
                              and separations take the nearest image
                              (physics3D.cpp:160-164, 206-210, 445-449) */
    float domain_w, domain_h, domain_d; /* myVar.domainSize3D (parameters.h:99) */
    int32_t sph;           /* isSPHEnabled: SPH3D::pcisphSolver after gravity
                              (globalLogic.cpp:1707-1709) and the speed clamp of
                              integrateStart3D (physics3D.cpp:761-781) */
    int32_t temperature;   /* isTempEnabled: unsupported */
    int32_t compensated;   /* Kahan-compensated force accumulation (1 = on) */
    uint32_t leaf_capacity; /* 0: the reference's leaf rule (quadtree.cpp:120-121);
                               1..64: at most this many particles per leaf */
    uint32_t walk_block;   /* threads per workgroup of the walk: 0 (= 256), 64, 128, 256 */
    float sph_h;           /* SPH3D::radiusMultiplier (SPH3D.h:13) */
    float sph_cell;        /* NeighborSearch3DV2AVX2::cellSize (neighborSearch.h:981) */
    float sph_unit_mass;   /* myVar.mass */
    float viscosity;       /* myVar.viscosity */
    float cohesion_coef;   /* myVar.cohesionCoefficient */
    float delta;           /* myVar.delta */
    float stiff_mult;      /* myVar.stiffMultiplier */
    float sph_max_vel;     /* myVar.sphMaxVel */
    int32_t sph_ground;    /* myVar.sphGround: SPH3D::groundModeBoundary after
                              PCISPH (SPH3D.cpp:310-366), walls at +-domain/2 */
    float vertical_gravity;   /* myVar.verticalGravity (ground mode) */
    float boundary_friction;  /* myVar.boundaryFriction (ground mode) */
    uint32_t reserved[4];
} ge3_step_params;

/* Node3D as laid out by the reference (GE/include/Physics/quadtree.h:134-179)
 * in DFS pre-order; child[q] uses q&1 -> +x, q&2 -> +y, q&4 -> +z. */
typedef struct {
    float posx, posy, posz, size;
    float comx, comy, comz, mass;
    uint32_t start, end, next;
    uint32_t child[8];
} ge3_node;

void ge3_default_config(ge3_config *cfg);
/* reference defaults (parameters.h:99,114-118; globalLogic.cpp:1678) */
void ge3_default_params(ge3_step_params *p);
int ge3_create(ge3_ctx **out, const ge3_config *cfg);
int ge3_destroy(ge3_ctx *ctx);
const char *ge3_last_error(const ge3_ctx *ctx);
int ge3_sync(ge3_ctx *ctx);
int ge3_count(const ge3_ctx *ctx, uint32_t *n);

/* state transfer (replaces the host vector pParticles3D) */
int ge3_upload(ge3_ctx *ctx, uint32_t n, const ge3_particles *p);
int ge3_download(ge3_ctx *ctx, const ge3_particles *p);

/* glm::vec4 boundingBox3D() (globalLogic.cpp:1465-1477): {x, y, z, size} */
int ge3_bounding_box(ge3_ctx *ctx, float out_pos_size[4]);
/* Morton::computeMortonKeys3D (morton.cpp:94-113) against the given box */
int ge3_compute_morton_keys(ge3_ctx *ctx, const float pos_size[4]);
/* the 63-bit keys of the last ge3_compute_morton_keys, in pre-sort order */
int ge3_get_morton_keys(ge3_ctx *ctx, uint64_t *keys);
/* Morton::sortParticlesByMortonKey3D (morton.cpp:115-137), stable by (key, index) */
int ge3_sort_by_morton_key(ge3_ctx *ctx);
/* perm[i] = pre-sort index of the particle now in slot i */
int ge3_get_permutation(ge3_ctx *ctx, uint32_t *perm);
/* Octree root3D(pP, rP, pos, size) (quadtree.cpp:275-292) over the sorted particles */
int ge3_build_octree(ge3_ctx *ctx, const float pos_size[4]);
int ge3_get_node_count(ge3_ctx *ctx, uint32_t *count);
/* globalNodes3D in the reference layout */
int ge3_get_nodes(ge3_ctx *ctx, ge3_node *out, uint32_t cap);
/* acc = 0, then Physics3D::calculateForceFromGrid3D (physics3D.cpp:131-254),
 * or naiveGravity3D (:416-470) when naive or N < direct_below, as mode3D
 * dispatches (globalLogic.cpp:1667-1702) */
int ge3_compute_gravity(ge3_ctx *ctx, const ge3_step_params *p);
/* Physics3D::integrateStart3D (physics3D.cpp:740-804) with dt_start */
int ge3_integrate_start(ge3_ctx *ctx, const ge3_step_params *p);
/* Physics3D::integrateEnd3D (physics3D.cpp:1158-1168) with dt */
int ge3_integrate_end(ge3_ctx *ctx, const ge3_step_params *p);
/* `frames` whole frames of mode3D's physics (gravity, then SPH3D when
 * p->sph), fused on the device */
int ge3_step(ge3_ctx *ctx, const ge3_step_params *p, int32_t frames);
/* NeighborSearch3DV2AVX2::newGridAVX2 (neighborSearch.h:1017-1107) over pos */
int ge3_sph_new_grid(ge3_ctx *ctx, const ge3_step_params *p);
/* SPH3D::pcisphSolver (SPH3D.h:590-599): PCISPH (SPH3D.cpp:172-308) on the
 * current grid, then groundModeBoundary when p->sph_ground */
int ge3_sph_pcisph(ge3_ctx *ctx, const ge3_step_params *p);
/* SPH candidate sets of the current grid (queryNeighborsAVX2,
 * neighborSearch.h:1109-1173: the 27 cells around the query's cell, exact
 * cell match, no distance filter, self included) queried at pos (which = 0)
 * or predPos (which = 1): CSR offsets off[n + 1] and particle slots idx[]
 * (ascending per particle).  idx may be NULL to size it. */
int ge3_get_neighbors(ge3_ctx *ctx, int32_t which, int64_t *off, uint32_t *idx, int64_t cap, int64_t *total);
/* NeighborSearch3DV2AVX2::newGridAVX2 + neighborAmount (neighborSearch.h:
 * 1175-1193): candidate count per particle (self included) in device order;
 * pinned, drawn and dark-matter particles get 0 */
int ge3_neighbor_counts(ge3_ctx *ctx, const ge3_step_params *p, uint32_t *counts);

/* FP64 direct-sum accelerations (j != i, r^2 + softening) for nt target
 * slots of the current order, for force-error checks (no counterpart in the
 * reference) */
int ge3_direct_fp64(ge3_ctx *ctx, double G, double softening, uint32_t nt, const uint32_t *targets,
                    double *ax, double *ay, double *az);
/* interactions of the last walk: {particle-node + particle-particle, accepted nodes} */
int ge3_get_interaction_counts(ge3_ctx *ctx, uint64_t out[2]);

#ifdef __cplusplus
}
#endif
#endif
