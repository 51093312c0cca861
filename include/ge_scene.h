/*
 * ge_scene.h -- Galaxy Engine scene files on the device path: load a save
 * file into a 2D or 3D context and write the device state back in the same
 * format, so scenes saved by the application run on the GPU and GPU results
 * open in the application.
 *
 * The format is SaveSystem's (GE/src/UX/saveSystem.cpp:3-614,
 * GE/include/UX/saveSystem.h:30-611):
 *   - YAML text: one "Key: value" map per parameter (paramIO), documents
 *     separated by "---";
 *   - the line "---PARTICLE BINARY DATA---";
 *   - uint32 version (190; 180 is read the same way), uint32 globalId,
 *     globalShapeId, globalWallId, uint32 particle count;
 *   - per particle the ParticlePhysics(3D) and ParticleRendering(3D) fields
 *     written one by one in SaveSystem's order, packed (180 bytes in 2D, 208
 *     in 3D; saveSystem.cpp:309-396, saveSystem.h:228-342);
 *   - constraints, walls, shapes and lights (optics), kept here verbatim.
 * Whether the records are 2D or 3D comes from the YAML key is3DMode.
 *
 * Host code only (no device memory of its own); int status codes as in
 * ge_physics.h.  A scene is not thread-safe; use one per thread.
 */
#ifndef GE_SCENE_H
#define GE_SCENE_H

#include <stdint.h>

#include "ge_physics.h"
#include "ge_physics3d.h"
#include "ge_render.h"
#include "ge_thermal.h"
#include "ge_constraints.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ge_scene ge_scene;

typedef struct {
    int32_t is3d;             /* is3DMode */
    uint32_t version;         /* version of the binary section */
    uint32_t count;           /* particles */
    uint32_t record_bytes;    /* bytes per particle record (180 or 208) */
    uint32_t global_id, global_shape_id, global_wall_id;
    uint32_t reserved;
    uint64_t tail_bytes;      /* constraints, walls, shapes, lights (kept verbatim) */
} ge_scene_info;

/* SaveSystem::deserializeParticleSystem (saveSystem.h:121-199): parse a
 * file.  On failure *out is NULL and ge_scene_last_error(NULL) says why. */
int ge_scene_load(const char *path, ge_scene **out);
/* an empty scene (no particles, a YAML header with is3DMode, empty tail) */
int ge_scene_new(int32_t is3d, ge_scene **out);
void ge_scene_free(ge_scene *s);
const char *ge_scene_last_error(const ge_scene *s);
int ge_scene_get_info(const ge_scene *s, ge_scene_info *info);
/* the YAML value text of key (NULL when absent); valid until the scene changes */
const char *ge_scene_value(const ge_scene *s, const char *key);
/* sets (or appends) key: value in the YAML header */
int ge_scene_set_value(ge_scene *s, const char *key, const char *value);
/* the packed particle records (count * record_bytes bytes) */
int ge_scene_records(const ge_scene *s, const void **data, uint64_t *bytes);

/* The physics variables a frame reads, from the YAML keys over the
 * defaults of ge_default_params / ge3_default_params: G = 6.674e-11 *
 * GravityMultiplier (globalLogic.cpp:1135), dt = 0.045 * TimeMult
 * (globalLogic.cpp:1187), Theta, Softening, SPHEnabled, SPHRadiusMult,
 * SPHMass, SPHViscosity, SPHStiffness, SPHCohesion, SPHDelta, SPHMaxVel,
 * SPHGround, SPHGravity, LoopingSpace && !InfiniteDomain, Domain sizes,
 * TemperatureSimulation. */
int ge_scene_params(const ge_scene *s, ge_step_params *p);
int ge_scene_params3(const ge_scene *s, ge3_step_params *p);
/* the physics variables back into the YAML keys (the inverse mapping) */
int ge_scene_set_params(ge_scene *s, const ge_step_params *p);
int ge_scene_set_params3(ge_scene *s, const ge3_step_params *p);

/* the temperature constants from AmbientTemperature, AmbientHeatRate and
 * HeatConductivityMultiplier (saveSystem.cpp:103-106) over the defaults of
 * ge_default_thermal_params, and back */
int ge_scene_thermal_params(const ge_scene *s, ge_thermal_params *t);
int ge_scene_set_thermal_params(ge_scene *s, const ge_thermal_params *t);

/* the constraint solver's variables from ParticleConstraints,
 * UnbreakableConstraints, ConstraintsStiffMultiplier and
 * ConstraintsResistMultiplier (saveSystem.cpp:198-207) over the defaults of
 * ge_default_constraint_params (*enabled gets ParticleConstraints), and back */
int ge_scene_constraint_params(const ge_scene *s, ge_constraint_params *c, int32_t *enabled);
int ge_scene_set_constraint_params(ge_scene *s, const ge_constraint_params *c, int32_t enabled);

/* particles into a context (ge_upload / ge3_upload): pos, vel, acc, prevVel,
 * mass, id, the SPH fields when the context has them, the flags isPinned,
 * isSPH, isBeingDrawn, isDarkMatter, uniqueColor, isSolid, isHotPoint,
 * hasSolidified (GE_FLAG_*), color, pColor, sColor, size, neighbors when it
 * keeps render fields (ge_render.h), temp, sphLabel, ke, prevKe when it
 * keeps temperature fields (ge_thermal.h), and the 2D file's constraints
 * (the start of the tail) when it holds constraints (ge_constraints.h) */
int ge_scene_upload(const ge_scene *s, ge_ctx *ctx);
int ge_scene_upload3(const ge_scene *s, ge3_ctx *ctx);
/* the context's particles back into records, in the device order: the
 * device fields from the context (the render and temperature fields and the
 * constraints too when it keeps them),
 * every other field from the record of the same id (a new id gets the
 * reference's defaults for a spawned particle, particle.h:35-63, 100-130) */
int ge_scene_download(ge_scene *s, ge_ctx *ctx);
int ge_scene_download3(ge_scene *s, ge3_ctx *ctx);
/* SaveSystem::saveSystem's file layout (saveSystem.cpp:285-611) */
int ge_scene_save(const ge_scene *s, const char *path);

#ifdef __cplusplus
}
#endif
#endif /* GE_SCENE_H */
