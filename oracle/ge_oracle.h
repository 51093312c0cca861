/*
 * ge_oracle.h -- CPU restatement of Galaxy Engine's 2D per-frame physics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path and the timed CPU baseline ("port").  Only tests/, the smoke entry and
 * the bench's cpu_baseline leg load it.  The product library never links it.
 *
 * Every function restates the reference algorithm in plain C with the same
 * floating-point expression order.  It is compiled with -ffp-contract=off so
 * that no a*b+c is fused: the reference's build would fuse some of them
 * depending on compiler and flags, so one mode is pinned here and in the HIP
 * kernels.  Citations are path:line under the reference tree, with
 * GE/ = GalaxyEngine/.
 *
 * Pins: the golden KATs and invariant, closed-form and brute-force checks of
 * tests/test_oracle.py.  The reference cannot be compiled here (its pch.h
 * needs glm), so these parts are parity unpinned beyond invariants:
 * geo_sph_pcisph's doubled pair sums, the dens = 0 viscosity of the first
 * frame, the predDens / restDens equation of state and the double pow kernel
 * constants (pinned by momentum conservation, the isolated-particle density
 * and the kernel closed forms only).
 *
 * Differences from the reference that are deliberate and documented:
 *  - the Morton sort is stable by (key, original index); the reference uses
 *    std::sort, whose order inside equal-key runs is unspecified
 *    (GE/src/Physics/morton.cpp:51);
 *  - the BH leaf body follows the scalar variant (GE/src/Physics/physics.cpp:
 *    91-205).  The AVX2 variant's 8-wide leaf body only runs for leaves with at
 *    least 8 particles, which the leaf rule only allows below size 1e-6;
 *  - the direct sum uses the exact 1/sqrtf of naiveGravity
 *    (physics.cpp:392-434), not the vendor-specific approximate rsqrt of
 *    naiveGravityAVX2 (physics.cpp:436-540);
 *  - SPH neighbour candidates are enumerated cell by cell in particle-index
 *    order, not in the reference's hash-bucket order (summation order only);
 *  - the SPH cell size is a parameter (the reference fixes it at 3.0,
 *    GE/include/Particles/neighborSearch.h:604).
 */
#ifndef GE_ORACLE_H
#define GE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* per-particle flag bits (the reference keeps them in ParticleRendering,
 * GE/include/Particles/particle.h:75-140) */
#define GEO_PINNED 1u
#define GEO_SPH 2u
#define GEO_DRAWN 4u
#define GEO_DARK 8u

typedef struct {
    float posx, posy;   /* lower-left corner (Node::pos) */
    float comx, comy;   /* Node::centerOfMass */
    float size;         /* Node::size (side length) */
    float mass;         /* Node::gridMass */
    uint32_t start, end;
    uint32_t next;      /* number of descendants (pre-order skip) */
    uint32_t child[4];  /* child node index per quadrant q (q&1 -> +x, q&2 -> +y), or UINT32_MAX */
} geo_node;

typedef struct {
    int32_t n;
    float *x, *y, *vx, *vy, *ax, *ay, *pvx, *pvy, *m;
    uint32_t *flags, *id;
    /* SPH fields (may be NULL when SPH is off) */
    float *sph_mass, *rest_dens, *stiff, *visc, *cohesion, *dens;
    float *press, *press_tmp, *pred_dens, *ppx, *ppy, *pfx, *pfy;
    /* temperature fields (may be NULL when temperature is off): temp,
     * sphLabel, ke, prevKe; isHotPoint / hasSolidified ride in flags */
    float *temp;
    uint32_t *label;
    float *ke, *prev_ke;
} geo_state;

/* SPHMaterial (GE/include/Physics/materialsSPH.h:3-47), the fields the
 * physics reads */
typedef struct {
    uint32_t id;
    float mass_mult, rest_dens, stiff, visc, cohesion;
    float hot_point, hot_rest_dens, hot_mass_mult, hot_stiff, hot_visc, hot_cohesion;
    float cold_point, cold_rest_dens, cold_mass_mult, cold_stiff, cold_visc, cold_cohesion;
    float heat_conductivity;
    float constraint_resistance, constraint_plastic_point, constraint_plastic_point_mult, constraint_stiffness;
    int32_t is_plastic;
} geo_material;

#define GEO_HOT_POINT 64u
#define GEO_SOLIDIFIED 128u

typedef struct {
    float G, theta, softening;
    float dt_start;        /* timeFactor used by integrateStart (previous frame's) */
    float dt;              /* timeFactor used by the rest of the step */
    int32_t naive;         /* myVar.naive: always direct sum */
    int32_t direct_below;  /* direct sum when N < this (reference: 5000) */
    int32_t sph;           /* myVar.isSPHEnabled */
    float sph_h;           /* SPH::radiusMultiplier */
    float sph_cell;        /* neighbour-grid cell size */
    float sph_unit_mass;   /* myVar.mass */
    float viscosity, cohesion_coef, delta, stiff_mult, sph_max_vel;
    int32_t sph_search;       /* 0: NeighborSearchV2AVX2 semantics; 1: NeighborSearchV2 */
    float sph_search_radius;  /* NeighborSearchV2::searchRadius (3.5) */
    int32_t periodic;         /* isPeriodicBoundaryEnabled && !infiniteDomain */
    float domain_w, domain_h; /* myVar.domainSize */
    int32_t sph_ground;       /* myVar.sphGround: SPH::groundModeBoundary after PCISPH */
    float vertical_gravity;   /* myVar.verticalGravity */
    float boundary_friction;  /* myVar.boundaryFriction */
    int32_t temperature;      /* myVar.isTempEnabled */
    float heat_conductivity;  /* myVar.globalHeatConductivity (0.045) */
    float ambient_temp;       /* myVar.ambientTemp (274) */
    float ambient_heat_rate;  /* myVar.globalAmbientHeatRate (1) */
    const geo_material *materials; /* SPHMaterials::idToMaterial */
    int32_t nmaterials;
} geo_params;

/* domain of the periodic wrap: width and height, or 0 for no wrap */
typedef struct { float w, h; } geo_domain;

/* bounding box: GE/src/globalLogic.cpp:1083-1099 -> {x, y, size} */
void geo_bbox(int32_t n, const float *x, const float *y, float out[3]);
/* Morton: GE/src/Physics/morton.cpp:3-38 */
uint64_t geo_scale_to_grid(float pos, float lo, float hi);
uint64_t geo_spread_bits(uint64_t v);
uint64_t geo_morton2d(uint64_t ix, uint64_t iy);
void geo_morton_keys(int32_t n, const float *x, const float *y, const float bb[3], uint64_t *keys);
/* stable (key, index) sort -> perm[i] = source index of sorted slot i */
void geo_sort_perm(int32_t n, const uint64_t *keys, uint32_t *perm);
/* quadtree: GE/src/Physics/quadtree.cpp:5-107; returns node count or -1 if cap exceeded */
int32_t geo_build_tree(int32_t n, const float *x, const float *y, const float *m,
                       const float bb[3], geo_node *nodes, int32_t cap);
/* the same build with leaf rule "count <= leaf_cap || size <= 1e-6" when
 * leaf_cap > 0 (autotuned trees; leaf_cap 0 is geo_build_tree) */
int32_t geo_build_tree_cap(int32_t n, const float *x, const float *y, const float *m,
                           const float bb[3], uint32_t leaf_cap, geo_node *nodes, int32_t cap);
/* BH accelerations: GE/src/Physics/physics.cpp:91-205 */
void geo_bh_accel(int32_t n, const float *x, const float *y, const float *m,
                  const geo_node *nodes, int32_t nn, float G, float theta, float eps,
                  float *ax, float *ay, uint64_t *interactions);
/* the same with the periodic wrap of every separation (physics.cpp:117-120,
 * 169-172) when dom->w > 0 */
void geo_bh_accel_dom(int32_t n, const float *x, const float *y, const float *m,
                      const geo_node *nodes, int32_t nn, float G, float theta, float eps,
                      const geo_domain *dom, float *ax, float *ay, uint64_t *interactions);
/* direct sum: GE/src/Physics/physics.cpp:392-434 (epsilon NOT squared) */
void geo_direct_accel(int32_t n, const float *x, const float *y, const float *m,
                      float G, float eps, float *ax, float *ay);
void geo_direct_accel_dom(int32_t n, const float *x, const float *y, const float *m,
                          float G, float eps, const geo_domain *dom, float *ax, float *ay);
/* SPH::groundModeBoundary, GE/src/Physics/SPH.cpp:283-322 */
void geo_sph_ground(geo_state *s, float radius, float w, float h, float vertical_gravity, float friction);
/* FP64 direct sum with Plummer r^2+eps^2, j != i, for nt targets */
void geo_direct_fp64(int32_t n, const float *x, const float *y, const float *m,
                     double G, double eps, int32_t nt, const uint32_t *targets,
                     double *ax, double *ay);
/* integrator: GE/src/Physics/physics.cpp:1069-1125 */
void geo_integrate_start(geo_state *s, float dt, int32_t sph, float sph_max_vel);
/* with the periodic position wrap (physics.cpp:1104-1110) when dom->w > 0 */
void geo_integrate_start_dom(geo_state *s, float dt, int32_t sph, float sph_max_vel, const geo_domain *dom);
void geo_integrate_end(geo_state *s, float dt);
/* SPH neighbour candidates (AVX2 grid semantics, neighborSearch.h:636-777):
 * all particles whose cell of (x,y) is within one cell (3x3) of cell(qx,qy).
 * Writes CSR: off[n+1], idx[] (each list sorted ascending). Returns total or
 * -1 when cap is too small. */
int64_t geo_sph_candidates(int32_t n, const float *x, const float *y,
                           int32_t nq, const float *qx, const float *qy, float cell,
                           int64_t *off, uint32_t *idx, int64_t cap);
/* The same with the search mode: mode 1 follows NeighborSearchV2 (the
 * reference's fallback without AVX2, neighborSearch.h:439-599): cell =
 * (int)(p * (1/cell)) (truncation), ceil(radius / cell) cells around the
 * query cell, and |p_j - q|^2 <= radius^2 with p_j the grid position. */
int64_t geo_sph_candidates_mode(int32_t n, const float *x, const float *y,
                                int32_t nq, const float *qx, const float *qy, float cell,
                                int32_t mode, float radius, int64_t *off, uint32_t *idx, int64_t cap);
/* NeighborSearchV2::neighborAmount (neighborSearch.h:581-598): candidates
 * (self included) of every particle over all particles at pos; pinned,
 * drawn and dark-matter particles get 0 (the reference leaves them as they
 * were). */
void geo_neighbor_amount(int32_t n, const float *x, const float *y, const uint32_t *flags, float cell,
                         float radius, uint32_t *counts);
/* PCISPH (one iteration): GE/src/Physics/SPH.cpp:98-281 */
int32_t geo_sph_pcisph(geo_state *s, const geo_params *p);
/* SPH kernel constants computed as the reference does (double pow, float PI) */
void geo_sph_constants(float h, float out[3]); /* {W volume, spiky scale, laplacian scale} */
/* compiler and flags this library was built with (reported with CPU timings) */
const char *geo_build_info(void);
/* one full frame: GE/src/globalLogic.cpp:1155-1372 for the 2D path.
 * Permutes every state array into Morton order.  Returns 0 or negative. */
int32_t geo_step(geo_state *s, const geo_params *p, int32_t nsteps);

/* ---- constraints ---- */
/* ParticleConstraint (GE/include/Physics/constraint.h:3-14) */
typedef struct {
    uint32_t id1, id2;
    float rest_length, original_length, stiffness, resistance, displacement, plasticity_point;
    uint8_t is_broken, is_plastic, pad[2];
} geo_constraint;

typedef struct {
    float stiffness_mult, resistance_mult; /* globalConstraintStiffnessMult, globalConstraintResistance */
    int32_t unbreakable;                   /* unbreakableConstraints */
    int32_t forces;                        /* timeFactor > 0 && gridExists */
    float damping, correction_ratio;       /* globalConstraintDamping, stiffCorrectionRatio */
    int32_t periodic;
    float domain_w, domain_h;
} geo_cparams;

/* Physics::constraints (GE/src/Physics/physics.cpp:768-950) run serially:
 * removal (order kept), *frame_count += 1 when any remain, the triangle rule
 * when it is a multiple of 5, then 15 substeps visiting the constraints in
 * the order of ascending (color, index).  The serial loop is one of the
 * outcomes of the reference's parallel loop.  `color` travels with the
 * constraints through the removal.  Returns the remaining count. */
int32_t geo_constraints(geo_state *s, geo_constraint *cons, uint32_t *color, int32_t m, const geo_cparams *p,
                        const geo_material *mats, int32_t nmat, uint32_t *frame_count);

/* Physics::createConstraints' automatic branch (GE/src/Physics/physics.cpp:
 * 654-766 with no key, selection or drawing flag): every particle i with
 * hasSolidified (GEO_SOLIDIFIED), in index order, clears the flag and, for
 * each neighbour candidate j != i of its position (the search of `mode`,
 * `cell`, `radius`, as QueryNeighbors::queryNeighbors returns them) with
 * |p_j - p_i|^2 <= 12, id_i < id_j and not (both materials known, cold
 * point of i zero and of j non-zero), appends {id_i, id_j, d, d,
 * stiffness, resistance, 0, plasticity} (materials of i and j: the
 * stiffness of i, the means of resistance and plastic point; otherwise 60,
 * 0.6, 0.6).  Candidates are visited cell by cell, dx outer and dy inner as
 * the reference loops, ascending index within a cell.  At most cap are
 * written; returns the number appended, or -1 when cap was too small. */
int32_t geo_create_constraints(geo_state *s, const geo_material *mats, int32_t nmat, int32_t mode, float cell,
                               float radius, geo_constraint *out, int32_t cap);

/* ---- merger ---- */
typedef struct {
    float density_radius;    /* NeighborSearch::densityRadius */
    float cell_size;         /* NeighborSearch::cellSize */
    float texture_half_size; /* particleTextureHalfSize (16) */
    float size_multiplier;   /* particleSizeMultiplier (0.6) */
} geo_mparams;

/* NeighborSearch::UpdateNeighbors (GE/include/Particles/neighborSearch.h:48-200;
 * cells filled in index order, as a serial run fills them) and
 * Physics::mergerSolver (GE/src/Physics/physics.cpp:954-1067) with
 * totalRadius = size * halfSize * sizeMultiplier (globalLogic.cpp:1225-1227).
 * size, prev_size, neighbors are the render fields and move with the
 * particles through the swap-and-pop deletions, as every state field does.
 * Returns the particle count left. */
int32_t geo_merge(geo_state *s, float *size, float *prev_size, int32_t *neighbors, const geo_mparams *p);

/* ---- temperature (isTempEnabled) ---- */
/* Node::gridTemp (GE/include/Physics/quadtree.h:53-100): leaves sum their
 * particles' temp in order, internal nodes their children's in the order
 * q = 0, 2, 1, 3 */
void geo_node_temps(const geo_node *nodes, int32_t nn, const float *temp, float *gtemp);
/* the heat transfer of calculateForceFromGridAVX2 (GE/src/Physics/physics.cpp:
 * 206-360; the x86-64 build's walk): every accepted node moves temp[i] toward
 * its average temperature, in walk order; opened leaves transfer nothing */
void geo_walk_heat(int32_t n, const float *x, const float *y, const geo_node *nodes, int32_t nn,
                   const float *gtemp, float theta, float eps, float conductivity, float dt,
                   const geo_domain *dom, float *temp);
/* Physics::temperatureCalculation (GE/src/Physics/physics.cpp:567-652) */
void geo_temperature(geo_state *s, const geo_material *mats, int32_t nmat, float ambient, float rate, float dt);

#ifdef __cplusplus
}
#endif
#endif
