/*
 * ge_thermal.h -- the temperature coupling of the 2D frame on the device
 * (isTempEnabled): the heat transfer that rides on the Barnes-Hut walk
 * (calculateForceFromGridAVX2, GE/src/Physics/physics.cpp:206-360, the walk
 * the x86-64 build takes), the node temperatures it reads (Node::gridTemp,
 * GE/include/Physics/quadtree.h:53-100) and Physics::temperatureCalculation
 * (GE/src/Physics/physics.cpp:567-652) with the SPH materials
 * (GE/include/Physics/materialsSPH.h).  Same conventions as ge_physics.h.
 *
 * A context created with enable_temperature (and enable_sph: the materials
 * write the SPH fields) keeps, per particle and moved by every sort, temp,
 * sphLabel, ke and prevKe; isHotPoint and hasSolidified ride in the flags
 * (GE_FLAG_HOT_POINT, GE_FLAG_SOLIDIFIED).  ge_step with
 * ge_step_params.temperature = 1 then runs the frame as updateScene() does:
 * the walk moves heat between each particle and the nodes it accepts (the
 * direct sum below direct_below moves none, as naiveGravityAVX2), PCISPH
 * runs, temperatureCalculation updates temperatures and material phases,
 * then integrateEnd.  The walk's heat transfer follows the reference's
 * monopole walk (accuracy 0).
 */
#ifndef GE_THERMAL_H
#define GE_THERMAL_H

#include <stdint.h>

#include "ge_physics.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SPHMaterial (materialsSPH.h:3-47): the fields the physics reads */
typedef struct {
    uint32_t id;                 /* the sphLabel that selects it */
    float mass_mult, rest_dens, stiff, visc, cohesion;
    float hot_point, hot_rest_dens, hot_mass_mult, hot_stiff, hot_visc, hot_cohesion;
    float cold_point, cold_rest_dens, cold_mass_mult, cold_stiff, cold_visc, cold_cohesion;
    float heat_conductivity;
    float constraint_resistance, constraint_plastic_point, constraint_plastic_point_mult, constraint_stiffness;
    int32_t is_plastic;
} ge_material;

#define GE_MAX_MATERIALS 64

/* The UpdateVariables fields of the coupling (GE/include/parameters.h:120-124). */
typedef struct {
    float heat_conductivity;  /* globalHeatConductivity (0.045) */
    float ambient_temp;       /* ambientTemp (274) */
    float ambient_heat_rate;  /* globalAmbientHeatRate (1) */
    float base_mass;          /* UpdateVariables::particleBaseMass (8.5e9) */
    uint32_t reserved[4];
} ge_thermal_params;

/* host arrays of the temperature fields (NULL arrays are skipped) */
typedef struct {
    float *temp;        /* temp */
    uint32_t *label;    /* ParticleRendering::sphLabel: the material id, 0 for none */
    float *ke, *prev_ke;
} ge_thermal_fields;

void ge_default_thermal_params(ge_thermal_params *t);
/* *on = 1 when the context keeps temperature fields (enable_temperature) */
int ge_thermal_enabled(const ge_ctx *ctx, int32_t *on);
/* the seven materials SPHMaterials::Init registers (materialsSPH.h:49-300):
 * water, rock, iron, sand, soil, mud, rubber (ids 1-7); *count gets 7 */
int ge_default_materials(ge_material *out, uint32_t cap, uint32_t *count);
/* the coupling's constants and the material table (at most GE_MAX_MATERIALS;
 * a label with no material takes the reference's defaults, conductivity
 * 0.05 and no phases) */
int ge_set_thermal(ge_ctx *ctx, const ge_thermal_params *t, const ge_material *materials, uint32_t count);
/* the fields of the current particles, in the context's current order;
 * particles that enter through ge_upload start at temp 288, label 0 */
int ge_thermal_upload(ge_ctx *ctx, const ge_thermal_fields *f);
int ge_thermal_download(ge_ctx *ctx, const ge_thermal_fields *f);
/* Physics::temperatureCalculation alone, on the current state */
int ge_temperature_calculation(ge_ctx *ctx, const ge_step_params *p);
/* Node::gridTemp of the current tree's nodes in the order of ge_get_nodes
 * (computed from the current temperatures) */
int ge_get_node_temps(ge_ctx *ctx, float *out, uint32_t cap);

#ifdef __cplusplus
}
#endif
#endif /* GE_THERMAL_H */
