/*
 * ge_constraints.h -- the particle constraints of the 2D frame on the device:
 * Physics::constraints (GE/src/Physics/physics.cpp:768-950): removal of
 * constraints whose particles are gone or that broke, the triangle rule
 * every fifth frame, and 15 substeps of spring, damping and position
 * correction with the breaking and plastic rules.  Same conventions as
 * ge_physics.h.  GE/ = GalaxyEngine/.
 *
 * A context created with max_constraints > 0 holds up to that many
 * constraints in HBM.  The reference updates every constraint of a substep
 * in parallel with atomic adds, so its result depends on the thread
 * schedule.  Here the constraints are coloured so that no two of one colour
 * share a particle, and each substep runs the colours one after another:
 * the result is the serial loop in (colour, index) order, one of the
 * outcomes the reference's parallel loop allows, and it is deterministic.
 * The colouring is redone only when constraints are added.
 *
 * Materials (ge_thermal.h) decide breaking: a constraint breaks or yields
 * only when both particles' sphLabel name a material
 * (SPHMaterials::idToMaterial), as in the reference; a context without
 * temperature fields has no labels, so its constraints never break.
 */
#ifndef GE_CONSTRAINTS_H
#define GE_CONSTRAINTS_H

#include <stdint.h>

#include "ge_physics.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ParticleConstraint (GE/include/Physics/constraint.h:3-14), the same 36-byte
 * layout the save files hold (SaveSystem writes the vector raw,
 * saveSystem.cpp:421-429) */
typedef struct {
    uint32_t id1, id2;
    float rest_length, original_length;
    float stiffness, resistance;
    float displacement, plasticity_point;
    uint8_t is_broken, is_plastic, pad[2];
} ge_constraint;

/* the UpdateVariables fields of the solver (GE/include/parameters.h:221-230;
 * physics.h:25-27 for the two constants) */
typedef struct {
    float stiffness_mult;    /* globalConstraintStiffnessMult (1) */
    float resistance_mult;   /* globalConstraintResistance (1) */
    int32_t unbreakable;     /* unbreakableConstraints */
    int32_t forces;          /* timeFactor > 0 && gridExists: apply the spring forces */
    float damping;           /* globalConstraintDamping (0.001) */
    float correction_ratio;  /* stiffCorrectionRatio (0.013333) */
    uint32_t reserved[2];
} ge_constraint_params;

void ge_default_constraint_params(ge_constraint_params *c);
/* *cap = max_constraints of the context */
int ge_constraints_capacity(const ge_ctx *ctx, uint32_t *cap);
int ge_set_constraint_params(ge_ctx *ctx, const ge_constraint_params *c);
/* replaces the context's constraints (particleConstraints) */
int ge_constraints_upload(ge_ctx *ctx, const ge_constraint *c, uint32_t count);
/* the constraints in their current order; *count gets their number, at
 * most cap are written */
int ge_constraints_download(ge_ctx *ctx, ge_constraint *out, uint32_t cap, uint32_t *count);
/* deleteAllConstraints */
int ge_constraints_clear(ge_ctx *ctx);
/* Physics::constraints once, on the current state (ge_step with
 * ge_step_params.constraints = 1 runs it after PCISPH, as updateScene does) */
int ge_constraints_solve(ge_ctx *ctx, const ge_step_params *p);
/* Physics::createConstraints (GE/src/Physics/physics.cpp:654-766), the
 * branch that runs without user input: every particle with hasSolidified
 * (GE_FLAG_SOLIDIFIED, set by temperatureCalculation) consumes the flag and
 * links to each neighbour candidate j of its position (the grid and search
 * of p: sph_cell, sph_search, sph_search_radius) with |p_j - p_i|^2 <= 12,
 * id_i < id_j and not (both materials known, cold point of i zero and of j
 * non-zero): rest and original length the distance, the stiffness of i's
 * material and the mean resistance and plastic point of the two (60, 0.6,
 * 0.6 without materials).  The constraints are appended in slot order of i,
 * then the reference's cell order (dx outer, dy inner), ascending slot
 * within a cell.  ge_step with constraints = 1 on a context with
 * temperature fields runs this every frame after the tree, as updateScene
 * does (globalLogic.cpp:1299-1300).  Needs enable_sph and max_constraints;
 * *created gets the number appended.  GE_E_CAPACITY when they did not all
 * fit (the excess is dropped). */
int ge_create_constraints(ge_ctx *ctx, const ge_step_params *p, uint32_t *created);
/* the order the substeps visit constraints (indices into the current
 * constraints, colour by colour) and the colour boundaries; *ncolors gets
 * their number (for checking the solve against a serial loop) */
int ge_constraints_order(ge_ctx *ctx, uint32_t *order, uint32_t cap, uint32_t *color_start, uint32_t color_cap,
                         uint32_t *ncolors);

#ifdef __cplusplus
}
#endif
#endif /* GE_CONSTRAINTS_H */
