/*
 * ge_oracle3d.h -- CPU restatement of Galaxy Engine's 3D physics step
 * (mode3D(), GE/src/globalLogic.cpp:1552-1722: gravity and SPH3D).
 *
 * TEST INFRASTRUCTURE ONLY, like ge_oracle.h: the checker for the HIP 3D
 * path and nothing else.  Same conventions: the reference's expression order,
 * compiled with -ffp-contract=off; citations under GE/ = GalaxyEngine/.
 *
 * Pins (tests/test_oracle3d.py): the 3D Morton keys against an independent
 * magic-mask spread and hand-worked known answers; the bbox against numpy;
 * the octree against brute-force invariants (every particle inside its leaf
 * cube up to the reference's partition-point rule, child ranges tiling the
 * parent, pre-order `next` counts, masses equal to the float sums of the
 * leaves in the reference order); the walk at theta = 0 against the direct
 * sum (i != j) and at any theta against the interaction count bound; the
 * direct sum against an FP64 closed form; the periodic wrap against a
 * hand-placed pair.  SPH3D: the candidate sets against an O(N^2) brute force
 * over cell coordinates, the kernel constants against their closed forms
 * (the W normalisation integral is 1), pair forces against antisymmetry
 * (momentum conserved without pinned particles) and a two-particle closed
 * form, the ground walls against hand-placed particles.  Parity unpinned:
 * the order of the reference's atomic SPH sums (nondeterministic there).
 *
 * Deliberate differences from the reference, as in 2D: the Morton sort is
 * stable by (key, index) (the reference's std::sort, morton.cpp:115-124, leaves
 * the order of equal keys unspecified); the BH leaf body and the direct sum
 * follow the scalar variants (physics3D.cpp:131-254, 416-470), not the AVX2
 * ones.
 */
#ifndef GE_ORACLE3D_H
#define GE_ORACLE3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float posx, posy, posz;   /* Node3D::pos (lower corner) */
    float comx, comy, comz;   /* Node3D::centerOfMass */
    float size, mass;         /* Node3D::size, gridMass */
    uint32_t start, end;
    uint32_t next;            /* descendants (pre-order skip) */
    uint32_t child[8];        /* per octant q (q&1 +x, q&2 +y, q&4 +z), or UINT32_MAX */
} geo3_node;

typedef struct {
    int32_t n;
    float *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *pvx, *pvy, *pvz, *m;
    uint32_t *flags, *id;
    /* SPH3D fields (may be NULL when SPH is off) */
    float *sph_mass, *rest_dens, *stiff, *visc, *cohesion, *dens;
    float *press, *press_tmp, *pred_dens;
    float *ppx, *ppy, *ppz, *pfx, *pfy, *pfz;
} geo3_state;

/* the cube of the periodic wrap, centred on the origin (domainSize3D); w = 0: none */
typedef struct { float w, h, d; } geo3_domain;

typedef struct {
    float G, theta, softening;
    float dt_start;          /* timeFactor seen by integrateStart3D */
    float dt;                /* timeFactor of the rest of the frame */
    int32_t naive;           /* myVar.naive */
    int32_t direct_below;    /* direct sum when N < this (reference: 5000) */
    int32_t sph_clamp;       /* isSPHEnabled: the speed clamp of integrateStart3D */
    float sph_max_vel;
    int32_t periodic;        /* isPeriodicBoundaryEnabled && !infiniteDomain */
    float domain_w, domain_h, domain_d;
    uint32_t leaf_cap;       /* 0: the reference's leaf rule */
    int32_t sph;             /* isSPHEnabled: SPH3D::pcisphSolver after gravity */
    float sph_h;             /* SPH3D::radiusMultiplier */
    float sph_cell;          /* NeighborSearch3DV2AVX2::cellSize */
    float sph_unit_mass;     /* myVar.mass */
    float viscosity, cohesion_coef, delta, stiff_mult;
    int32_t sph_ground;      /* myVar.sphGround */
    float vertical_gravity, boundary_friction;
} geo3_params;

/* boundingBox3D, GE/src/globalLogic.cpp:1465-1477 -> {x, y, z, size} */
void geo3_bbox(int32_t n, const float *x, const float *y, const float *z, float out[4]);
/* Morton::scaleToGrid3D / spreadBits3D / morton3D, morton.cpp:67-92 */
uint64_t geo3_scale_to_grid(float pos, float lo, float hi);
uint64_t geo3_spread_bits(uint64_t v);
uint64_t geo3_morton(uint64_t ix, uint64_t iy, uint64_t iz);
/* Morton::computeMortonKeys3D, morton.cpp:94-113 */
void geo3_morton_keys(int32_t n, const float *x, const float *y, const float *z, const float bb[4],
                      uint64_t *keys);
/* Octree / Node3D, GE/src/Physics/quadtree.cpp:109-292; returns the node count
 * or -1.  leaf_cap > 0: leaf when count <= leaf_cap (size floor kept). */
int32_t geo3_build_tree(int32_t n, const float *x, const float *y, const float *z, const float *m,
                        const float bb[4], uint32_t leaf_cap, geo3_node *nodes, int32_t cap);
/* Physics3D::calculateForceFromGrid3D, GE/src/Physics/physics3D.cpp:131-254 */
void geo3_bh_accel(int32_t n, const float *x, const float *y, const float *z, const float *m,
                   const geo3_node *nodes, int32_t nn, float G, float theta, float eps,
                   const geo3_domain *dom, float *ax, float *ay, float *az, uint64_t *interactions);
/* Physics3D::naiveGravity3D, physics3D.cpp:416-470 (softening added unsquared) */
void geo3_direct_accel(int32_t n, const float *x, const float *y, const float *z, const float *m,
                       float G, float eps, const geo3_domain *dom, float *ax, float *ay, float *az);
/* FP64 direct sum, j != i, with the reference's 3D softening r^2 + eps */
void geo3_direct_fp64(int32_t n, const float *x, const float *y, const float *z, const float *m,
                      double G, double eps, int32_t nt, const uint32_t *targets,
                      double *ax, double *ay, double *az);
/* integrateStart3D / integrateEnd3D, physics3D.cpp:740-805, 1158-1168 */
void geo3_integrate_start(geo3_state *s, float dt, int32_t sph_clamp, float sph_max_vel,
                          const geo3_domain *dom);
void geo3_integrate_end(geo3_state *s, float dt);
/* SPH3D kernel constants (SPH3D.h:27-53): {W volume 2 pi h^5 / 15, spiky
 * scale -45 / (pi h^6), laplacian scale 45 / (pi h^6)}, computed as the 2D
 * ones are (double pow, float PI, rounded to float) */
void geo3_sph_constants(float h, float out[3]);
/* queryNeighborsAVX2 candidate sets (neighborSearch.h:1109-1173) of the
 * grid over (x, y, z) for nq query points: CSR off[nq + 1], idx ascending per
 * query.  Returns the total, or -1 when it exceeds cap (off still filled). */
int64_t geo3_sph_candidates(int32_t n, const float *x, const float *y, const float *z, int32_t nq,
                            const float *qx, const float *qy, const float *qz, float cell,
                            int64_t *off, uint32_t *idx, int64_t cap);
/* SPH3D::PCISPH, GE/src/Physics/SPH3D.cpp:98-308 (one iteration, maxIter 1),
 * on a grid over the current positions */
int32_t geo3_sph_pcisph(geo3_state *s, const geo3_params *p);
/* SPH3D::groundModeBoundary, GE/src/Physics/SPH3D.cpp:310-366 */
void geo3_sph_ground(geo3_state *s, float radius, float w, float h, float d, float vertical_gravity,
                     float friction);
/* frames of mode3D (globalLogic.cpp:1580-1719): integrate, bbox, keys, sort
 * (every array permuted), octree, forces, SPH3D, kick.  0 or < 0. */
int32_t geo3_step(geo3_state *s, const geo3_params *p, int32_t nsteps);

#ifdef __cplusplus
}
#endif
#endif
