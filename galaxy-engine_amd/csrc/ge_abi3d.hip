// ge_abi3d.hip -- the extern "C" entry points of include/ge_physics3d.h:
// context lifecycle, device buffers, state transfer, the stage calls and the
// frame driver that orders the 3D stage kernels as mode3D() does
// (GE/src/globalLogic.cpp:1580-1719: gravity, then SPH3D).
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "ge_internal3d.hpp"

using namespace ge;

namespace {

template <typename T>
hipError_t dalloc(T *&p, size_t count) {
    p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * count);
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

#define GE3_CHECK(c, expr)                                                     \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            (c)->err = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return GE_E_HIP;                                                   \
        }                                                                      \
    } while (0)

int fail(ge3_ctx *c, int code, const char *msg) {
    if (c) c->err = msg;
    return code;
}

void free_all(ge3_ctx *c) {
    Buffers3 &b = c->b;
    for (State3 &s : b.st) {
        float **fs[] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.ax, &s.ay, &s.az, &s.pvx, &s.pvy, &s.pvz, &s.m};
        for (float **f : fs) dfree(*f);
        dfree(s.flags);
        dfree(s.id);
        dfree(s.color); dfree(s.pcolor); dfree(s.scolor); dfree(s.size); dfree(s.nbrs); dfree(s.prev_size);
    }
    dfree(b.pxyzm); dfree(b.keys); dfree(b.sk_a); dfree(b.sk_b); dfree(b.sv_a); dfree(b.sv_b); dfree(b.perm);
    dfree(b.hist); dfree(b.partials); dfree(b.bb); dfree(b.counters); dfree(b.inter);
    TopNodes3 &T = b.top;
    dfree(T.px); dfree(T.py); dfree(T.pz); dfree(T.size); dfree(T.start); dfree(T.end); dfree(T.child); dfree(T.exact);
    dfree(T.kind); dfree(T.pending2); dfree(T.offset); dfree(T.broot); dfree(T.blist); dfree(T.blist_off);
    dfree(T.blist_cnt); dfree(T.depth);
    dfree(b.start_count); dfree(b.start_level); dfree(b.scan_part); dfree(b.kstart); dfree(b.kdirty);
    dfree(b.nodes); dfree(b.node_geom); dfree(b.node_mask); dfree(b.diag);
    for (State3 &s : b.st) {
        float **fs[] = {&s.sph_mass, &s.rest_dens, &s.stiff, &s.visc, &s.cohesion, &s.dens, &s.press, &s.press_tmp,
                        &s.pred_dens, &s.ppx, &s.ppy, &s.ppz, &s.pfx, &s.pfy, &s.pfz};
        for (float **f : fs) dfree(*f);
    }
    for (Grid3 *g : {&b.grid_pos, &b.grid_pred}) {
        dfree(g->keys); dfree(g->keys_alt); dfree(g->first); dfree(g->ord); dfree(g->cell);
        dfree(g->pk0); dfree(g->pk1); dfree(g->pk2);
    }
    dfree(b.sphf); dfree(b.sph_pp); dfree(b.sph_pd);
    for (int k = 0; k < 2; ++k) {
        dfree(b.nbr_idx[k]);
        dfree(b.nbr_cnt[k]);
    }
}

// SPH3D state: 15 fields per state buffer, two slot grids, pair-force sums
int alloc_sph(ge3_ctx *c) {
    Buffers3 &b = c->b;
    const uint32_t cap = c->cap;
    uint32_t bits = 10;
    while (bits < 24 && (1ull << bits) < 2ull * cap) ++bits;  // about two slots per particle
    b.slot_bits = bits;
    for (State3 &s : b.st) {
        float **fs[] = {&s.sph_mass, &s.rest_dens, &s.stiff, &s.visc, &s.cohesion, &s.dens, &s.press, &s.press_tmp,
                        &s.pred_dens, &s.ppx, &s.ppy, &s.ppz, &s.pfx, &s.pfy, &s.pfz};
        for (float **f : fs) GE3_CHECK(c, dalloc(*f, cap));
    }
    for (Grid3 *g : {&b.grid_pos, &b.grid_pred}) {
        GE3_CHECK(c, dalloc(g->keys, cap));
        GE3_CHECK(c, dalloc(g->keys_alt, cap));
        GE3_CHECK(c, dalloc(g->first, (1ull << bits) + 1));
        GE3_CHECK(c, dalloc(g->ord, cap));
        GE3_CHECK(c, dalloc(g->cell, cap));
        GE3_CHECK(c, dalloc(g->pk0, cap));
        GE3_CHECK(c, dalloc(g->pk1, cap));
    }
    GE3_CHECK(c, dalloc(b.grid_pos.pk2, cap));
    GE3_CHECK(c, dalloc(b.sphf, cap));
    GE3_CHECK(c, dalloc(b.sph_pp, cap));
    GE3_CHECK(c, dalloc(b.sph_pd, cap));
    // about 34 kept candidates per particle at the SPH3D test densities; a
    // denser query falls back to rescanning its candidates
    b.nbr_cap = 96;
    for (int k = 0; k < 2; ++k) {
        GE3_CHECK(c, dalloc(b.nbr_idx[k], (size_t)cap * b.nbr_cap));
        GE3_CHECK(c, dalloc(b.nbr_cnt[k], cap));
    }
    return GE_OK;
}

int alloc_all(ge3_ctx *c) {
    Buffers3 &b = c->b;
    const uint32_t cap = c->cap;
    for (State3 &s : b.st) {
        float **fs[] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.ax, &s.ay, &s.az, &s.pvx, &s.pvy, &s.pvz, &s.m};
        for (float **f : fs) GE3_CHECK(c, dalloc(*f, cap));
        GE3_CHECK(c, dalloc(s.flags, cap));
        GE3_CHECK(c, dalloc(s.id, cap));
        if (c->render) {
            GE3_CHECK(c, dalloc(s.color, cap));
            GE3_CHECK(c, dalloc(s.pcolor, cap));
            GE3_CHECK(c, dalloc(s.scolor, cap));
            GE3_CHECK(c, dalloc(s.size, cap));
            GE3_CHECK(c, dalloc(s.nbrs, cap));
            GE3_CHECK(c, dalloc(s.prev_size, cap));
        }
    }
    GE3_CHECK(c, dalloc(b.pxyzm, cap));
    GE3_CHECK(c, dalloc(b.keys, cap));
    GE3_CHECK(c, dalloc(b.sk_a, cap));
    GE3_CHECK(c, dalloc(b.sk_b, cap));
    GE3_CHECK(c, dalloc(b.sv_a, cap));
    GE3_CHECK(c, dalloc(b.sv_b, cap));
    GE3_CHECK(c, dalloc(b.perm, cap));
    GE3_CHECK(c, dalloc(b.hist, radix_hist_entries(cap)));
    GE3_CHECK(c, hipMemset(b.hist, 0, sizeof(uint32_t) * radix_hist_entries(cap)));
    GE3_CHECK(c, dalloc(b.partials, 8ull * blocks_for(cap, kBBoxChunk3)));
    GE3_CHECK(c, dalloc(b.bb, 12));
    GE3_CHECK(c, dalloc(b.counters, kCntSlots));
    GE3_CHECK(c, hipMemset(b.counters, 0, sizeof(uint32_t) * kCntSlots));
    GE3_CHECK(c, dalloc(b.inter, 2));
    TopNodes3 &T = b.top;
    T.cap = 2u * cap + 1024u;
    GE3_CHECK(c, dalloc(T.px, T.cap)); GE3_CHECK(c, dalloc(T.py, T.cap)); GE3_CHECK(c, dalloc(T.pz, T.cap));
    GE3_CHECK(c, dalloc(T.size, T.cap)); GE3_CHECK(c, dalloc(T.start, T.cap)); GE3_CHECK(c, dalloc(T.end, T.cap));
    GE3_CHECK(c, dalloc(T.child, 8ull * T.cap)); GE3_CHECK(c, dalloc(T.kind, T.cap));
    GE3_CHECK(c, dalloc(T.pending2, T.cap)); GE3_CHECK(c, dalloc(T.offset, T.cap));
    GE3_CHECK(c, dalloc(T.broot, T.cap)); GE3_CHECK(c, dalloc(T.blist, T.cap));
    GE3_CHECK(c, dalloc(T.blist_off, T.cap)); GE3_CHECK(c, dalloc(T.blist_cnt, T.cap));
    GE3_CHECK(c, dalloc(T.depth, T.cap));
    GE3_CHECK(c, dalloc(T.exact, T.cap));
    for (uint32_t *u : {T.start, T.end, T.kind, T.pending2, T.offset, T.broot, T.blist_cnt, T.depth, T.exact})
        GE3_CHECK(c, hipMemset(u, 0, sizeof(uint32_t) * T.cap));
    GE3_CHECK(c, hipMemset(T.child, 0xFF, sizeof(uint32_t) * 8ull * T.cap));
    GE3_CHECK(c, dalloc(b.start_count, (size_t)cap + 1));
    GE3_CHECK(c, dalloc(b.start_level, (size_t)cap + 1));
    GE3_CHECK(c, dalloc(b.scan_part, (size_t)cap / 1024 + 64));
    GE3_CHECK(c, dalloc(b.kstart, kKeyStartEntries3));
    GE3_CHECK(c, dalloc(b.kdirty, kKeyDirtyEntries3));
    // Octree::root reserves 8N (quadtree.cpp:279); measured octrees hold
    // about 1.3N nodes, so 4N with an overflow flag
    b.node_cap = 4u * cap + 1024u;
    GE3_CHECK(c, dalloc(b.nodes, b.node_cap));
    GE3_CHECK(c, dalloc(b.node_geom, b.node_cap));
    GE3_CHECK(c, dalloc(b.node_mask, b.node_cap));
    return GE_OK;
}

int check_params(ge3_ctx *c, const ge3_step_params *p) {
    if (!p) return fail(c, GE_E_INVALID, "params is NULL");
    if (p->sph && !c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    if (p->sph && !(p->sph_h > 0.0f && p->sph_cell > 0.0f)) return fail(c, GE_E_INVALID, "sph_h and sph_cell must be > 0");
    if (p->sph && p->sph_ground && !(p->domain_w > 0.0f && p->domain_h > 0.0f && p->domain_d > 0.0f))
        return fail(c, GE_E_INVALID, "ground mode needs domain_w, domain_h, domain_d > 0");
    if (p->temperature) return fail(c, GE_E_UNSUPPORTED, "temperature transfer is not implemented");
    if (p->periodic && !(p->domain_w > 0.0f && p->domain_h > 0.0f && p->domain_d > 0.0f))
        return fail(c, GE_E_INVALID, "periodic mode needs domain_w, domain_h, domain_d > 0");
    if (p->leaf_capacity > 64) return fail(c, GE_E_INVALID, "leaf_capacity must be 0 (reference rule) or 1..64");
    if (p->walk_block != 0 && p->walk_block != 64 && p->walk_block != 128 && p->walk_block != 256)
        return fail(c, GE_E_INVALID, "walk_block must be 0, 64, 128 or 256");
    return GE_OK;
}

int check_device_errors(ge3_ctx *c) {
    uint32_t e = 0;
    GE3_CHECK(c, hipMemcpyAsync(&e, c->b.counters + kCntError, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    if (e) {
        char buf[200];
        describe_device_errors(e, buf, sizeof buf);
        c->err = buf;
        uint32_t z = 0;
        (void)hipMemcpy(c->b.counters + kCntError, &z, sizeof z, hipMemcpyHostToDevice);
        return GE_E_DEVICE;
    }
    return GE_OK;
}

Domain3 domain_of(const ge3_step_params *p) {
    Domain3 d;
    if (p->periodic) {
        d.w = p->domain_w;
        d.h = p->domain_h;
        d.d = p->domain_d;
    }
    return d;
}

bool use_direct(const ge3_ctx *c, const ge3_step_params *p) {
    // globalLogic.cpp:1677-1701: naive flag, or N < 5000 on either code path
    return p->naive || (int64_t)c->n < (int64_t)p->direct_below;
}

// One gravity frame of mode3D (globalLogic.cpp:1580-1719).
void frame(ge3_ctx *c, const ge3_step_params *p, float dt_start, bool count) {
    c->leaf_cap = p->leaf_capacity;
    const bool sph = p->sph != 0;
    if (dt_start != 0.0f) {
        // integrateStart3D + boundingBox3D, computeMortonKeys3D, the sort, the octree
        launch3_integrate_bbox(c, dt_start, 1, domain_of(p), sph ? 1 : 0, p->sph_max_vel);
        launch3_morton_keys(c, c->b.bb);
        launch3_sort_gather(c);
        c->have_grid = false;
        if (!p->naive) {
            launch_octree_build(c, c->b.bb, true);
            c->have_tree = true;
        }
    }
    if (p->dt == 0.0f) return;
    // newGridAVX2 over the sorted positions (globalLogic.cpp:1647)
    if (sph) sph3_new_grid(c, p);
    // integrateEnd3D is fused into the last kernel that writes acc: the force
    // kernel, or the pressure pass when SPH runs (unless the ground walls
    // follow it)
    const float end_kick = p->dt * 0.5f;
    const float grav_kick = sph ? 0.0f : end_kick;
    if (p->G == 0.0f) {
        launch3_zero_acc(c);
        if (!sph) launch3_integrate_end(c, p->dt);
    } else if (use_direct(c, p) || !c->have_tree) {
        launch3_direct(c, p, grav_kick);
    } else {
        launch3_bh(c, p, grav_kick, count);
    }
    if (sph) {  // sph3D.pcisphSolver (globalLogic.cpp:1707-1709)
        sph3_pcisph(c, p, p->sph_ground ? 0.0f : p->dt);
        if (p->sph_ground) {
            sph3_ground(c, p);
            launch3_integrate_end(c, p->dt);
        }
    }
}

int put_fields(ge3_ctx *c, uint32_t n, const ge3_particles *p) {
    State3 &s = c->b.st[0];
    const size_t fb = sizeof(float) * n;
    const float *src[13] = {p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay, p->az, p->pvx, p->pvy, p->pvz,
                            p->mass};
    float *dst[13] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az, s.pvx, s.pvy, s.pvz, s.m};
    for (int k = 0; k < 13; ++k) {
        if (src[k]) GE3_CHECK(c, hipMemcpyAsync(dst[k], src[k], fb, hipMemcpyHostToDevice, c->stream));
        else GE3_CHECK(c, hipMemsetAsync(dst[k], 0, fb, c->stream));
    }
    if (p->flags) GE3_CHECK(c, hipMemcpyAsync(s.flags, p->flags, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    else GE3_CHECK(c, hipMemsetAsync(s.flags, 0, sizeof(uint32_t) * n, c->stream));
    std::vector<uint32_t> ids;
    if (p->id) {
        GE3_CHECK(c, hipMemcpyAsync(s.id, p->id, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    } else {
        ids.resize(n);
        for (uint32_t i = 0; i < n; ++i) ids[i] = i;
        GE3_CHECK(c, hipMemcpyAsync(s.id, ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    }
    if (c->sph) {
        const float *ssrc[15] = {p->sph_mass, p->rest_dens, p->stiff, p->visc, p->cohesion, p->dens, p->press,
                                 p->press_tmp, p->pred_dens, p->ppx, p->ppy, p->ppz, p->pfx, p->pfy, p->pfz};
        float *sdst[15] = {s.sph_mass, s.rest_dens, s.stiff, s.visc, s.cohesion, s.dens, s.press, s.press_tmp,
                           s.pred_dens, s.ppx, s.ppy, s.ppz, s.pfx, s.pfy, s.pfz};
        std::vector<float> ones;
        for (int k = 0; k < 15; ++k) {
            if (ssrc[k]) {
                GE3_CHECK(c, hipMemcpyAsync(sdst[k], ssrc[k], fb, hipMemcpyHostToDevice, c->stream));
            } else if (k == 0) {  // sph_mass defaults to 1
                ones.assign(n, 1.0f);
                GE3_CHECK(c, hipMemcpyAsync(sdst[k], ones.data(), fb, hipMemcpyHostToDevice, c->stream));
            } else {
                GE3_CHECK(c, hipMemsetAsync(sdst[k], 0, fb, c->stream));
            }
        }
        GE3_CHECK(c, hipStreamSynchronize(c->stream));
    }
    // ParticleRendering3D() defaults: white, size 1, no neighbours
    if (c->render) launch_render_fill(c->stream, s.color, s.pcolor, s.scolor, s.size, s.prev_size, s.nbrs, 0, n,
                                      0xFFFFFFFFu, 1.0f);
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

}  // namespace

extern "C" {

void ge3_default_config(ge3_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->capacity = 1u << 20;
    cfg->device = 0;
}

void ge3_default_params(ge3_step_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->G = 6.674e-11f;      // parameters.h:114
    p->theta = 0.8f;        // parameters.h:117
    p->softening = 1.5f;    // parameters.h:116
    p->dt_start = 1.0f;     // timeFactor (parameters.h:133)
    p->dt = 1.0f;
    p->direct_below = 5000; // globalLogic.cpp:1678
    p->domain_w = 1300.0f;  // domainSize3D (parameters.h:99)
    p->domain_h = 1300.0f;
    p->domain_d = 1300.0f;
    p->compensated = 1;
    p->sph_h = 3.0f;             // SPH3D::radiusMultiplier (SPH3D.h:13)
    p->sph_cell = 3.0f;          // NeighborSearch3DV2AVX2::cellSize (neighborSearch.h:981)
    p->sph_unit_mass = 0.03f;    // parameters.h:311
    p->viscosity = 0.3f;
    p->cohesion_coef = 1.0f;
    p->delta = 19000.0f;
    p->stiff_mult = 1.0f;
    p->sph_max_vel = 250.0f;     // parameters.h:118
    p->vertical_gravity = 3.0f;  // parameters.h:316
    p->boundary_friction = 0.0f; // parameters.h:361
}

int ge3_create(ge3_ctx **out, const ge3_config *cfg) {
    if (!out || !cfg) return GE_E_INVALID;
    *out = nullptr;
    if (cfg->capacity == 0 || cfg->capacity > kMaxParticles) return GE_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GE_E_NODEVICE;
    if (cfg->device < 0 || cfg->device >= ndev) return GE_E_NODEVICE;
    ge3_ctx *c = new (std::nothrow) ge3_ctx();
    if (!c) return GE_E_NOMEM;
    c->device = cfg->device;
    c->cap = cfg->capacity;
    if (hipSetDevice(c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GE_E_HIP;
    }
    c->sph = cfg->enable_sph != 0;
    c->render = cfg->enable_render != 0;
    int ra = alloc_all(c);
    if (ra == GE_OK && c->sph) ra = alloc_sph(c);
    if (int r = ra) {
        free_all(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return r == GE_E_HIP ? GE_E_NOMEM : r;
    }
    *out = c;
    return GE_OK;
}

int ge3_destroy(ge3_ctx *c) {
    if (!c) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GE_OK;
}

const char *ge3_last_error(const ge3_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ge3_sync(ge3_ctx *c) {
    if (!c) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return check_device_errors(c);
}

int ge3_count(const ge3_ctx *c, uint32_t *n) {
    if (!c || !n) return GE_E_INVALID;
    *n = c->n;
    return GE_OK;
}

int ge3_upload(ge3_ctx *c, uint32_t n, const ge3_particles *p) {
    if (!c || !p) return GE_E_INVALID;
    if (n > c->cap) return fail(c, GE_E_CAPACITY, "particle count above capacity");
    if (n > 0 && (!p->x || !p->y || !p->z || !p->mass)) return fail(c, GE_E_INVALID, "x, y, z and mass are required");
    (void)hipSetDevice(c->device);
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    c->n = n;
    c->b.cur = 0;
    c->have_tree = false;
    c->have_grid = false;
    if (n == 0) return GE_OK;
    if (int r = put_fields(c, n, p)) return r;
    launch3_pack(c);
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge3_download(ge3_ctx *c, const ge3_particles *p) {
    if (!c || !p) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    const uint32_t n = c->n;
    if (n > 0) {
        const State3 &s = c->b.st[c->b.cur];
        const size_t fb = sizeof(float) * n;
        float *dst[13] = {p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay, p->az, p->pvx, p->pvy, p->pvz, p->mass};
        const float *src[13] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az, s.pvx, s.pvy, s.pvz, s.m};
        for (int k = 0; k < 13; ++k)
            if (dst[k]) GE3_CHECK(c, hipMemcpyAsync(dst[k], src[k], fb, hipMemcpyDeviceToHost, c->stream));
        if (p->flags)
            GE3_CHECK(c, hipMemcpyAsync(p->flags, s.flags, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
        if (p->id) GE3_CHECK(c, hipMemcpyAsync(p->id, s.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
        if (c->sph) {
            float *sdst[15] = {p->sph_mass, p->rest_dens, p->stiff, p->visc, p->cohesion, p->dens, p->press,
                               p->press_tmp, p->pred_dens, p->ppx, p->ppy, p->ppz, p->pfx, p->pfy, p->pfz};
            const float *ssrc[15] = {s.sph_mass, s.rest_dens, s.stiff, s.visc, s.cohesion, s.dens, s.press,
                                     s.press_tmp, s.pred_dens, s.ppx, s.ppy, s.ppz, s.pfx, s.pfy, s.pfz};
            for (int k = 0; k < 15; ++k)
                if (sdst[k]) GE3_CHECK(c, hipMemcpyAsync(sdst[k], ssrc[k], fb, hipMemcpyDeviceToHost, c->stream));
        }
    }
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return check_device_errors(c);
}

int ge3_bounding_box(ge3_ctx *c, float out[4]) {
    if (!c || !out) return GE_E_INVALID;
    if (c->n == 0) return fail(c, GE_E_INVALID, "no particles");
    (void)hipSetDevice(c->device);
    launch3_integrate_bbox(c, 0.0f, 0, Domain3{});
    float h[4];
    GE3_CHECK(c, hipMemcpyAsync(h, c->b.bb, sizeof h, hipMemcpyDeviceToHost, c->stream));
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) out[k] = h[k];
    return GE_OK;
}

int ge3_compute_morton_keys(ge3_ctx *c, const float bb[4]) {
    if (!c || !bb) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    GE3_CHECK(c, hipMemcpyAsync(c->b.bb, bb, 4 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launch3_morton_keys(c, c->b.bb);
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge3_get_morton_keys(ge3_ctx *c, uint64_t *keys) {
    if (!c || !keys) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    GE3_CHECK(c, hipMemcpyAsync(keys, c->b.keys, sizeof(uint64_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge3_sort_by_morton_key(ge3_ctx *c) {
    if (!c) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch3_sort_gather(c);
    c->have_tree = false;
    c->have_grid = false;
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge3_get_permutation(ge3_ctx *c, uint32_t *perm) {
    if (!c || !perm) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    GE3_CHECK(c, hipMemcpyAsync(perm, c->b.perm, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge3_build_octree(ge3_ctx *c, const float bb[4]) {
    if (!c || !bb) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    GE3_CHECK(c, hipMemcpyAsync(c->b.bb, bb, 4 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    c->leaf_cap = 0;  // the reference's leaf rule
    launch_octree_build(c, c->b.bb);
    c->have_tree = true;
    return ge3_sync(c);
}

int ge3_get_node_count(ge3_ctx *c, uint32_t *count) {
    if (!c || !count) return GE_E_INVALID;
    *count = 0;
    if (!c->have_tree || c->n == 0) return GE_OK;
    GE3_CHECK(c, hipMemcpyAsync(count, c->b.counters + kCntNodeTotal, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                c->stream));
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return check_device_errors(c);
}

int ge3_get_nodes(ge3_ctx *c, ge3_node *out, uint32_t cap) {
    if (!c || !out) return GE_E_INVALID;
    uint32_t nn = 0;
    if (int r = ge3_get_node_count(c, &nn)) return r;
    if (nn > cap) return fail(c, GE_E_CAPACITY, "node buffer too small");
    std::vector<NodeRec3> rec(nn);
    std::vector<float4> geo(nn);
    std::vector<uint32_t> mask(nn);
    if (nn) {
        GE3_CHECK(c, hipMemcpy(rec.data(), c->b.nodes, sizeof(NodeRec3) * nn, hipMemcpyDeviceToHost));
        GE3_CHECK(c, hipMemcpy(geo.data(), c->b.node_geom, sizeof(float4) * nn, hipMemcpyDeviceToHost));
        GE3_CHECK(c, hipMemcpy(mask.data(), c->b.node_mask, sizeof(uint32_t) * nn, hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < nn; ++i) {
        ge3_node &o = out[i];
        o.posx = geo[i].x; o.posy = geo[i].y; o.posz = geo[i].z; o.size = geo[i].w;
        o.comx = rec[i].comx; o.comy = rec[i].comy; o.comz = rec[i].comz; o.mass = rec[i].mass;
        o.start = rec[i].start; o.end = rec[i].end; o.next = rec[i].next;
        uint32_t ch = i + 1;
        for (int q = 0; q < 8; ++q) {
            if (mask[i] & (1u << q)) {
                o.child[q] = ch;
                ch += (ch < nn ? rec[ch].next : 0u) + 1;
            } else {
                o.child[q] = kNone;
            }
        }
    }
    return GE_OK;
}

int ge3_compute_gravity(ge3_ctx *c, const ge3_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    if (p->G == 0.0f) launch3_zero_acc(c);
    else if (use_direct(c, p) || !c->have_tree) launch3_direct(c, p, 0.0f);
    else launch3_bh(c, p, 0.0f, true);
    return ge3_sync(c);
}

int ge3_integrate_start(ge3_ctx *c, const ge3_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch3_integrate_bbox(c, p->dt_start, 1, domain_of(p), p->sph ? 1 : 0, p->sph_max_vel);
    launch3_pack(c);
    c->have_tree = false;
    c->have_grid = false;
    return ge3_sync(c);
}

int ge3_integrate_end(ge3_ctx *c, const ge3_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch3_integrate_end(c, p->dt);
    return ge3_sync(c);
}

int ge3_step(ge3_ctx *c, const ge3_step_params *p, int32_t frames) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (frames < 0) return fail(c, GE_E_INVALID, "frames < 0");
    if (c->n == 0 || frames == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    for (int32_t k = 0; k < frames; ++k) frame(c, p, k == 0 ? p->dt_start : p->dt, false);
    GE3_CHECK(c, hipGetLastError());
    return GE_OK;
}

int ge3_sph_new_grid(ge3_ctx *c, const ge3_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (!c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    if (!(p->sph_cell > 0.0f)) return fail(c, GE_E_INVALID, "sph_cell must be > 0");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    sph3_new_grid(c, p);
    return ge3_sync(c);
}

int ge3_sph_pcisph(ge3_ctx *c, const ge3_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (!c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    if (!(p->sph_h > 0.0f)) return fail(c, GE_E_INVALID, "sph_h must be > 0");
    if (c->n == 0) return GE_OK;
    if (!c->have_grid || c->grid_cell != p->sph_cell)
        return fail(c, GE_E_INVALID, "no neighbour grid for the current positions and sph_cell (ge3_sph_new_grid)");
    (void)hipSetDevice(c->device);
    sph3_pcisph(c, p, 0.0f);
    if (p->sph_ground) sph3_ground(c, p);
    return ge3_sync(c);
}

int ge3_get_neighbors(ge3_ctx *c, int32_t which, int64_t *off, uint32_t *idx, int64_t cap, int64_t *total) {
    if (!c || !off || !total || (which != 0 && which != 1)) return GE_E_INVALID;
    if (!c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    *total = 0;
    off[0] = 0;
    if (c->n == 0) return GE_OK;
    if (!c->have_grid) return fail(c, GE_E_INVALID, "no neighbour grid (ge3_sph_new_grid)");
    (void)hipSetDevice(c->device);
    if (int r = sph3_neighbors(c, which, off, idx, cap, total)) return r;
    return check_device_errors(c);
}

int ge3_neighbor_counts(ge3_ctx *c, const ge3_step_params *p, uint32_t *counts) {
    if (!c || !counts) return GE_E_INVALID;
    if (int r = ge3_sph_new_grid(c, p)) return r;
    if (c->n == 0) return GE_OK;
    if (int r = sph3_neighbor_counts(c, counts)) return r;
    return check_device_errors(c);
}

int ge3_direct_fp64(ge3_ctx *c, double G, double eps, uint32_t nt, const uint32_t *targets, double *ax, double *ay,
                    double *az) {
    if (!c || (nt && (!targets || !ax || !ay || !az))) return GE_E_INVALID;
    if (nt == 0) return GE_OK;
    for (uint32_t k = 0; k < nt; ++k)
        if (targets[k] >= c->n) return fail(c, GE_E_INVALID, "target slot out of range");
    (void)hipSetDevice(c->device);
    const size_t need = 3ull * nt + (nt + 1) / 2 + 1;  // results, then the target list
    if (c->b.diag_cap < need) {
        dfree(c->b.diag);
        c->b.diag_cap = 0;
        GE3_CHECK(c, dalloc(c->b.diag, need));
        c->b.diag_cap = need;
    }
    uint32_t *dt = reinterpret_cast<uint32_t *>(c->b.diag + 3ull * nt);
    GE3_CHECK(c, hipMemcpyAsync(dt, targets, sizeof(uint32_t) * nt, hipMemcpyHostToDevice, c->stream));
    launch3_direct_fp64(c, G, eps, nt, dt, c->b.diag);
    std::vector<double> h(3ull * nt);
    GE3_CHECK(c, hipMemcpyAsync(h.data(), c->b.diag, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < nt; ++k) {
        ax[k] = h[3ull * k];
        ay[k] = h[3ull * k + 1];
        az[k] = h[3ull * k + 2];
    }
    return GE_OK;
}

int ge3_get_interaction_counts(ge3_ctx *c, uint64_t out[2]) {
    if (!c || !out) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    GE3_CHECK(c, hipMemcpyAsync(out, c->b.inter, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    GE3_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

}  // extern "C"
