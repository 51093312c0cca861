/*
 * ge_merger.h -- the particle merger of the 2D frame on the device:
 * NeighborSearch::UpdateNeighbors (GE/include/Particles/neighborSearch.h:
 * 48-200) and Physics::mergerSolver (GE/src/Physics/physics.cpp:954-1067).
 * Same conventions as ge_physics.h.  GE/ = GalaxyEngine/.
 *
 * Needs a context with render fields (ge_render.h): totalRadius is
 * size * particleTextureHalfSize * particleSizeMultiplier
 * (globalLogic.cpp:1225-1227), the merger grows previousSize, and
 * UpdateNeighbors writes `neighbors`.  In updateScene() the merger runs
 * after the forces and before PCISPH (globalLogic.cpp:1355-1356); with the
 * staged entry points that is ge_compute_gravity, ge_merge, then the rest.
 * Particles merged away leave by the reference's swap-with-back deletions,
 * so the survivors' order is the reference's.  The result is the
 * reference's serial loop, with each cell's particles in index order.
 */
#ifndef GE_MERGER_H
#define GE_MERGER_H

#include <stdint.h>

#include "ge_physics.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float density_radius;    /* NeighborSearch::densityRadius: 10 once the merger runs (3.5 before) */
    float cell_size;         /* NeighborSearch::cellSize (see ge_merger_cell_size) */
    float texture_half_size; /* particleTextureHalfSize (16) */
    float size_multiplier;   /* particleSizeMultiplier (0.6) */
    uint32_t reserved[4];
} ge_merger_params;

void ge_default_merger_params(ge_merger_params *m);
/* the cell size updateScene() sets for the next UpdateNeighbors
 * (globalLogic.cpp:1272-1292): clamp(2 * max totalRadius of the particles
 * that are not dark matter, 2, 80) */
int ge_merger_cell_size(ge_ctx *ctx, const ge_merger_params *m, float *cell);
/* UpdateNeighbors + mergerSolver; *removed (may be NULL) gets the number of
 * particles merged away */
int ge_merge(ge_ctx *ctx, const ge_merger_params *m, uint32_t *removed);

#ifdef __cplusplus
}
#endif
#endif /* GE_MERGER_H */
