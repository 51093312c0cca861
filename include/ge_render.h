/*
 * ge_render.h -- the steps either side of the physics frame, kept on the
 * device: the galaxy spawner that fills a context, the post-step consumers
 * of the neighbour counts (particle size and colour), and the PLY export.
 * Same conventions as ge_physics.h (int status codes, the context owns the
 * device memory, calls ordered on its stream).  GE/ = GalaxyEngine/.
 *
 * Render fields.  A context created with enable_render (ge_config /
 * ge3_config) keeps, per particle and in the same (Morton) order as the
 * physics fields, the ParticleRendering fields these steps read and write
 * (GE/include/Particles/particle.h:75-140): color, pColor, sColor (RGBA8,
 * red in the low byte: r | g << 8 | b << 16 | a << 24), size, previousSize
 * and neighbors.  Every sort and system compaction moves them with the particle.
 * Particles that enter through ge_upload get ParticleRendering()'s defaults
 * (white colours, size and previousSize 1, neighbors 0); ge_render_upload
 * overrides them.
 * uniqueColor and isSolid ride in the flags (GE_FLAG_UNIQUE_COLOR,
 * GE_FLAG_SOLID).
 */
#ifndef GE_RENDER_H
#define GE_RENDER_H

#include <stdint.h>

#include "ge_physics.h"
#include "ge_physics3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* host arrays of the render fields (NULL arrays are skipped) */
typedef struct {
    uint32_t *color, *pcolor, *scolor; /* color, pColor, sColor (RGBA8) */
    float *size;                       /* size */
    int32_t *neighbors;                /* neighbors */
    float *previous_size;              /* previousSize (grown by the merger) */
} ge_render_fields;

/* *on = 1 when the context keeps render fields (created with enable_render) */
int ge_render_enabled(const ge_ctx *ctx, int32_t *on);
int ge3_render_enabled(const ge3_ctx *ctx, int32_t *on);
/* the fields of the current particles, in the context's current order */
int ge_render_upload(ge_ctx *ctx, const ge_render_fields *f);
int ge_render_download(ge_ctx *ctx, const ge_render_fields *f);
int ge3_render_upload(ge3_ctx *ctx, const ge_render_fields *f);
int ge3_render_download(ge3_ctx *ctx, const ge_render_fields *f);

/* The switches and constants of the post-step consumers:
 * DensitySize (GE/include/Particles/densitySize.h:5-80) and the colour
 * modes of ColorVisuals::particlesColorVisuals that read the neighbour
 * counts or the acceleration (GE/include/Particles/particleColorVisuals.h:
 * 8-53, 111-277).  The frame order is the reference's: neighbour counts
 * (globalLogic.cpp:1229-1246, 3D :1632-1649), sizeByDensity
 * (globalLogic.cpp:1397), then the colours (globalLogic.cpp:2150). */
typedef struct {
    /* neighbour counts (NeighborSearch*::neighborAmount); particles that
     * are pinned, being drawn or dark matter keep their previous count */
    int32_t count_neighbors; /* 1: count before the consumers run; 0: use the stored counts */
    int32_t search;          /* 0: NeighborSearchV2AVX2 / NeighborSearch3DV2AVX2 (cell 3, the
                                3 x 3 (x 3) cells, self included; the x86-64 build);
                                1: NeighborSearchV2 (2D only: cell 3, radius 3.5) */
    float cell;              /* cellSize (3) */
    float radius;            /* searchRadius of search 1 (3.5) */
    /* DensitySize */
    int32_t density_size;    /* isDensitySizeEnabled */
    int32_t force_size;      /* isForceSizeEnabled */
    float min_size, max_size; /* 0.17, 0.62 */
    float size_acc;          /* 22 */
    int32_t size_max_neighbors; /* 200 */
    float size_multiplier;   /* particleSizeMultiplier (1) */
    /* ColorVisuals */
    int32_t solid_color;     /* solidColor */
    int32_t density_color;   /* densityColor */
    int32_t force_color;     /* forceColor */
    uint32_t p_color, s_color; /* pColor {0, 40, 68, 100}, sColor {155, 80, 40, 75} (RGBA8) */
    int32_t color_max_neighbors; /* 200 */
    float min_color_acc, max_color_acc; /* 0, 40 */
    uint32_t reserved[4];
} ge_visual_params;

void ge_default_visual_params(ge_visual_params *v);
int ge_visuals(ge_ctx *ctx, const ge_visual_params *v);
int ge3_visuals(ge3_ctx *ctx, const ge_visual_params *v);

/* One PLY vertex of plyFileCreation (GE/src/globalLogic.cpp:146-190) /
 * plyFileCreation3D (:192-242): the position as written (2D: scaled into
 * [-10 * w / h, 10 * w / h] x [-10, 10] by the domain, z = 0; 3D: pos),
 * color.rgb and size. */
typedef struct {
    float x, y, z;
    uint8_t r, g, b, pad;
    float radius;
} ge_ply_vertex;

/* The vertices of the particles that are not dark matter, in the context's
 * order, compacted on the device.  *count receives the number of vertices;
 * at most cap are written (out may be NULL with cap 0 to ask for the count).
 * 2D reads domain_w, domain_h (myVar.domainSize). */
int ge_ply_vertices(ge_ctx *ctx, float domain_w, float domain_h, ge_ply_vertex *out, uint32_t cap, uint32_t *count);
int ge3_ply_vertices(ge3_ctx *ctx, ge_ply_vertex *out, uint32_t cap, uint32_t *count);
/* The ASCII PLY text of exportPly (globalLogic.cpp:244-394) for those
 * vertices: the header, then "%.6g %.6g %.6g %d %d %d %.6g\n" per vertex. */
int ge_ply_write(ge_ctx *ctx, float domain_w, float domain_h, const char *path);
int ge3_ply_write(ge3_ctx *ctx, const char *path);
/* the text of one vertex array (len receives its length; at most cap bytes
 * are written, text may be NULL with cap 0) */
int ge_ply_format(const ge_ply_vertex *v, uint32_t count, char *text, uint64_t cap, uint64_t *len);

/* The galaxy tool of ParticlesSpawning::particlesInitialConditions
 * (GE/src/Particles/particlesSpawning.cpp:130-244): an exponential disc of
 * visible particles on circular orbits and, optionally, a dark-matter halo
 * (GE/include/Particles/particlesSpawning.h:28-32 for the radii). */
typedef struct {
    float center_x, center_y;  /* myCamera.mouseWorldPos */
    float extra_vx, extra_vy;  /* slingshot.norm * slingshot.length * 0.3 */
    uint32_t visible;          /* 40000 * particleAmountMultiplier */
    uint32_t dark;             /* 12000 * DMAmountMultiplier, 0 without dark matter */
    float outer_radius;        /* outerRadius (200) */
    float scale_length;        /* scaleLength (90) */
    float outer_radius_dm;     /* outerRadiusDM (2000) */
    float radius_core_dm;      /* radiusCoreDM (3.5) */
    float mass_visible;        /* 8.5e9 (/ particleAmountMultiplier with massMultiplierEnabled) */
    float mass_dark;           /* 1.416e11 (/ DMAmountMultiplier with massMultiplierEnabled) */
    float mass_scatter;        /* massScatter (0.75) */
    uint32_t first_id;         /* globalId: ids first_id, first_id + 1, ... */
    uint32_t reserved[4];
} ge_galaxy_params;

void ge_default_galaxy_params(ge_galaxy_params *g);
/* Appends visible + dark particles after the context's current ones (as
 * emplace_back does) and returns the index of the first in *first (may be
 * NULL).  The draws come from `uniforms` when it is not NULL (host floats in
 * [0, 1], 3 per visible particle: radius, angle, mass; then 5 per dark one:
 * radius, angle, vx, vy, mass; a caller can replay the reference's
 * generator with them), otherwise from a counter-based generator on the
 * device: draw d of particle k of stream s (0 visible, 1 dark) is
 * (splitmix64(seed + (s << 60) + 8 k + d) >> 40) * 2^-24.  A single-system
 * context only; render fields as ParticleRendering(...) in the tool. */
int ge_spawn_galaxy(ge_ctx *ctx, const ge_galaxy_params *g, uint64_t seed, const float *uniforms, uint32_t *first);

#ifdef __cplusplus
}
#endif
#endif /* GE_RENDER_H */
