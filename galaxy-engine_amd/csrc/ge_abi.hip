// ge_abi.hip -- the extern "C" entry points of include/ge_physics.h:
// context lifecycle, device buffers, state transfer and the frame driver that
// orders the stage kernels as updateScene() does (globalLogic.cpp:1155-1372).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "ge_internal.hpp"

namespace ge {
uint32_t radix_hist_entries(uint32_t cap);
}

using namespace ge;

namespace {

thread_local uint64_t t_alloc_bytes = 0;  // bytes allocated by this thread's current ge_create

template <typename T>
hipError_t dalloc(T *&p, size_t count) {
    p = nullptr;
    if (count == 0) count = 1;
    t_alloc_bytes += sizeof(T) * count;
    return hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * count);
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int alloc_state(ge_ctx *c, State &s, uint32_t cap, bool sph) {
    float **fs[] = {&s.x, &s.y, &s.vx, &s.vy, &s.ax, &s.ay, &s.pvx, &s.pvy, &s.m};
    for (float **f : fs) GE_HIP_CHECK(c, dalloc(*f, cap));
    GE_HIP_CHECK(c, dalloc(s.flags, cap));
    GE_HIP_CHECK(c, dalloc(s.id, cap));
    if (sph) {
        float **gs[] = {&s.sph_mass, &s.rest_dens, &s.stiff, &s.visc, &s.cohesion, &s.dens, &s.press,
                        &s.press_tmp, &s.pred_dens, &s.ppx, &s.ppy, &s.pfx, &s.pfy};
        for (float **f : gs) GE_HIP_CHECK(c, dalloc(*f, cap));
    }
    if (c->thermal) {
        GE_HIP_CHECK(c, dalloc(s.temp, cap));
        GE_HIP_CHECK(c, dalloc(s.ke, cap));
        GE_HIP_CHECK(c, dalloc(s.prev_ke, cap));
        GE_HIP_CHECK(c, dalloc(s.label, cap));
    }
    if (c->render) {
        GE_HIP_CHECK(c, dalloc(s.color, cap));
        GE_HIP_CHECK(c, dalloc(s.pcolor, cap));
        GE_HIP_CHECK(c, dalloc(s.scolor, cap));
        GE_HIP_CHECK(c, dalloc(s.size, cap));
        GE_HIP_CHECK(c, dalloc(s.nbrs, cap));
        GE_HIP_CHECK(c, dalloc(s.prev_size, cap));
    }
    if (c->precise) {  // zero home forms restart from the first positions they meet
        GE_HIP_CHECK(c, dalloc(s.hx, cap));
        GE_HIP_CHECK(c, dalloc(s.hy, cap));
        GE_HIP_CHECK(c, dalloc(s.ox, cap));
        GE_HIP_CHECK(c, dalloc(s.oy, cap));
        GE_HIP_CHECK(c, hipMemset(s.hx, 0, sizeof(int32_t) * cap));
        GE_HIP_CHECK(c, hipMemset(s.hy, 0, sizeof(int32_t) * cap));
        GE_HIP_CHECK(c, hipMemset(s.ox, 0, sizeof(float) * cap));
        GE_HIP_CHECK(c, hipMemset(s.oy, 0, sizeof(float) * cap));
    }
    return GE_OK;
}

void free_state(State &s) {
    float **fs[] = {&s.x, &s.y, &s.vx, &s.vy, &s.ax, &s.ay, &s.pvx, &s.pvy, &s.m, &s.sph_mass, &s.rest_dens,
                    &s.stiff, &s.visc, &s.cohesion, &s.dens, &s.press, &s.press_tmp, &s.pred_dens, &s.ppx,
                    &s.ppy, &s.pfx, &s.pfy};
    for (float **f : fs) dfree(*f);
    dfree(s.flags);
    dfree(s.id);
    dfree(s.color); dfree(s.pcolor); dfree(s.scolor); dfree(s.size); dfree(s.nbrs); dfree(s.prev_size);
    dfree(s.temp); dfree(s.ke); dfree(s.prev_ke); dfree(s.label);
    dfree(s.hx); dfree(s.hy); dfree(s.ox); dfree(s.oy);
}

void free_all(ge_ctx *c) {
    if (c->diag) (void)hipFree(c->diag);
    c->diag = nullptr;
    c->diag_cap = 0;
    Buffers &b = c->b;
    free_state(b.st[0]);
    free_state(b.st[1]);
    dfree(b.pxym); dfree(b.pho); dfree(b.node_home); dfree(b.sort_a); dfree(b.sort_b); dfree(b.keys); dfree(b.perm); dfree(b.hist);
    dfree(b.partials); dfree(b.partials2); dfree(b.counters); dfree(b.interactions);
    dfree(b.sys_off); dfree(b.sys_of); dfree(b.sys_bb); dfree(b.chunk_lo); dfree(b.chunk_first); dfree(b.sys_src);
    dfree(b.sys_wave_off);
    dfree(b.ccreate);
    b.bb = nullptr;
    TopNodes &T = b.top;
    dfree(T.px); dfree(T.py); dfree(T.size); dfree(T.start); dfree(T.end); dfree(T.slot); dfree(T.exact); dfree(T.child);
    dfree(T.kind); dfree(T.pending2); dfree(T.offset); dfree(T.broot); dfree(T.depth);
    dfree(T.blist); dfree(T.blist_off); dfree(T.blist_cnt);
    dfree(b.start_count); dfree(b.start_level); dfree(b.scan_part); dfree(b.kstart); dfree(b.kdirty);
    merger_free(c);
    dfree(b.node_temp); dfree(b.node_tavg); dfree(b.node_parent); dfree(b.node_arrive); dfree(b.materials);
    dfree(b.cons); dfree(b.cons_tmp); dfree(b.cidx); dfree(b.cidx_tmp); dfree(b.ccolor); dfree(b.ccolor_tmp);
    dfree(b.ckeep); dfree(b.cblock); dfree(b.ccount); dfree(b.corder); dfree(b.cstart); dfree(b.cw_a);
    dfree(b.cw_b); dfree(b.chist);
    b.cframe = b.cleft = nullptr;
    dfree(b.nodes); dfree(b.node_geom); dfree(b.node_mp); dfree(b.mp_cnt); dfree(b.mp_wave_list); dfree(b.mp_big_list); dfree(b.mp_piece_owner); dfree(b.mp_piece_sum); dfree(b.sphf); dfree(b.sph_pp); dfree(b.sph_pd); dfree(b.sph_ppxy);
    for (Grid *g : {&b.grid_pos, &b.grid_pred}) {
        dfree(g->keys); dfree(g->keys_alt); dfree(g->first); dfree(g->ord); dfree(g->frame);
        dfree(g->pk0); dfree(g->pk1); dfree(g->pkf); dfree(g->xy);
    }
}

int alloc_all(ge_ctx *c) {
    Buffers &b = c->b;
    const uint32_t cap = c->cap;
    if (int r = alloc_state(c, b.st[0], cap, c->sph)) return r;
    if (int r = alloc_state(c, b.st[1], cap, c->sph)) return r;
    GE_HIP_CHECK(c, dalloc(b.pxym, cap));
    if (c->precise) GE_HIP_CHECK(c, dalloc(b.pho, cap));
    GE_HIP_CHECK(c, dalloc(b.sort_a, cap));
    GE_HIP_CHECK(c, dalloc(b.sort_b, cap));
    GE_HIP_CHECK(c, dalloc(b.keys, cap));
    GE_HIP_CHECK(c, dalloc(b.perm, cap));
    b.hist_cap = radix_hist_entries(cap);
    GE_HIP_CHECK(c, dalloc(b.hist, b.hist_cap));
    GE_HIP_CHECK(c, hipMemset(b.hist, 0, sizeof(uint32_t) * b.hist_cap));
    b.max_chunks = cap / kChunk + c->max_sys + 1;
    GE_HIP_CHECK(c, dalloc(b.partials, 4ull * b.max_chunks));
    GE_HIP_CHECK(c, dalloc(b.partials2, 4 * 4096));
    GE_HIP_CHECK(c, dalloc(b.counters, kCntSlots));
    GE_HIP_CHECK(c, hipMemset(b.counters, 0, sizeof(uint32_t) * kCntSlots));
    GE_HIP_CHECK(c, dalloc(b.sys_off, (size_t)c->max_sys + 1));
    GE_HIP_CHECK(c, dalloc(b.sys_of, cap));
    GE_HIP_CHECK(c, dalloc(b.sys_bb, 8ull * c->max_sys));
    GE_HIP_CHECK(c, dalloc(b.chunk_lo, (size_t)b.max_chunks + 1));
    GE_HIP_CHECK(c, dalloc(b.chunk_first, (size_t)c->max_sys + 1));
    GE_HIP_CHECK(c, dalloc(b.sys_src, (size_t)c->max_sys + 1));
    GE_HIP_CHECK(c, dalloc(b.sys_wave_off, (size_t)c->max_sys + 1));
    b.bb = b.sys_bb;  // system 0's box doubles as the single-scene box
    GE_HIP_CHECK(c, dalloc(b.interactions, 4));  // interactions, accepted nodes, wave visits, leaf trips
    TopNodes &T = b.top;
    T.cap = 2u * cap + 1024u;
    GE_HIP_CHECK(c, dalloc(T.px, T.cap)); GE_HIP_CHECK(c, dalloc(T.py, T.cap));
    GE_HIP_CHECK(c, dalloc(T.size, T.cap)); GE_HIP_CHECK(c, dalloc(T.start, T.cap));
    GE_HIP_CHECK(c, dalloc(T.end, T.cap)); GE_HIP_CHECK(c, dalloc(T.slot, T.cap));
    GE_HIP_CHECK(c, dalloc(T.exact, T.cap));
    GE_HIP_CHECK(c, dalloc(T.child, 4ull * T.cap)); GE_HIP_CHECK(c, dalloc(T.kind, T.cap));
    GE_HIP_CHECK(c, dalloc(T.pending2, T.cap));
    GE_HIP_CHECK(c, dalloc(T.offset, T.cap));
    GE_HIP_CHECK(c, dalloc(T.broot, T.cap));
    GE_HIP_CHECK(c, dalloc(T.blist, T.cap));
    GE_HIP_CHECK(c, dalloc(T.blist_off, T.cap));
    GE_HIP_CHECK(c, dalloc(T.blist_cnt, T.cap));
    GE_HIP_CHECK(c, dalloc(T.depth, T.cap));
    GE_HIP_CHECK(c, dalloc(b.start_count, (size_t)cap + 1));
    GE_HIP_CHECK(c, dalloc(b.start_level, (size_t)cap + 1));
    GE_HIP_CHECK(c, dalloc(b.scan_part, (size_t)cap / 1024 + 64));
    GE_HIP_CHECK(c, dalloc(b.kstart, kKeyStartEntries));
    GE_HIP_CHECK(c, dalloc(b.kdirty, kKeyDirtyEntries));
    // zeroed top nodes read as "internal" (never emitted) until a build writes them
    for (uint32_t *u : {T.start, T.end, T.slot, T.exact, T.kind, T.pending2, T.offset, T.broot, T.blist_cnt,
                        T.depth})
        GE_HIP_CHECK(c, hipMemset(u, 0, sizeof(uint32_t) * T.cap));
    GE_HIP_CHECK(c, hipMemset(T.child, 0xFF, sizeof(uint32_t) * 4ull * T.cap));
    // Quadtree::root reserves 4N nodes (quadtree.cpp:94)
    b.node_cap = 4u * cap + 1024u;
    GE_HIP_CHECK(c, dalloc(b.nodes, b.node_cap));
    GE_HIP_CHECK(c, dalloc(b.node_geom, b.node_cap));
    if (c->cons_cap) {
        const uint32_t mc = c->cons_cap;
        GE_HIP_CHECK(c, dalloc(b.cons, mc));
        GE_HIP_CHECK(c, dalloc(b.cons_tmp, mc));
        GE_HIP_CHECK(c, dalloc(b.cidx, mc));
        GE_HIP_CHECK(c, dalloc(b.cidx_tmp, mc));
        GE_HIP_CHECK(c, dalloc(b.ccolor, mc));
        GE_HIP_CHECK(c, dalloc(b.ccolor_tmp, mc));
        GE_HIP_CHECK(c, dalloc(b.ckeep, mc));
        GE_HIP_CHECK(c, dalloc(b.cblock, (size_t)mc / 256 + 2));
        GE_HIP_CHECK(c, dalloc(b.ccount, 4));
        GE_HIP_CHECK(c, hipMemset(b.ccount, 0, 4 * sizeof(uint32_t)));
        if (c->sph)  // createConstraints queries the SPH grid
            GE_HIP_CHECK(c, dalloc(b.ccreate, (size_t)cap + (cap + 1023) / 1024 + 4));
        b.cframe = b.ccount + 1;
        b.cleft = b.ccount + 2;
        GE_HIP_CHECK(c, dalloc(b.corder, mc));
        GE_HIP_CHECK(c, dalloc(b.cstart, 72));
        const uint64_t words = std::max<uint64_t>(2ull * mc, cap);
        GE_HIP_CHECK(c, dalloc(b.cw_a, words));
        GE_HIP_CHECK(c, dalloc(b.cw_b, words));
        GE_HIP_CHECK(c, dalloc(b.chist, radix_hist_entries((uint32_t)words)));
        GE_HIP_CHECK(c, hipMemset(b.chist, 0, sizeof(uint32_t) * radix_hist_entries((uint32_t)words)));
    }
    if (c->thermal) {
        GE_HIP_CHECK(c, dalloc(b.node_temp, b.node_cap));
        GE_HIP_CHECK(c, dalloc(b.node_tavg, b.node_cap));
        GE_HIP_CHECK(c, dalloc(b.node_parent, b.node_cap));
        GE_HIP_CHECK(c, dalloc(b.node_arrive, b.node_cap));
        GE_HIP_CHECK(c, dalloc(b.materials, GE_MAX_MATERIALS));
    }
    if (c->sph) {
        GE_HIP_CHECK(c, dalloc(b.sphf, 2ull * cap));
        GE_HIP_CHECK(c, dalloc(b.sph_pp, 4ull * cap));
        GE_HIP_CHECK(c, dalloc(b.sph_pd, 2ull * cap));
        GE_HIP_CHECK(c, dalloc(b.sph_ppxy, 2ull * cap));
        for (Grid *g : {&b.grid_pos, &b.grid_pred}) {
            // dense cell table: at least 4 cells per particle of capacity, at most 2^24
            uint64_t cells = 4ull * cap + 4096;
            if (cells < (1ull << 20)) cells = 1ull << 20;
            if (cells > (1ull << kCellBits) - 16) cells = (1ull << kCellBits) - 16;  // below the sentinel key
            g->cells_cap = (uint32_t)cells;
            GE_HIP_CHECK(c, dalloc(g->keys, cap));
            GE_HIP_CHECK(c, dalloc(g->keys_alt, cap));
            GE_HIP_CHECK(c, dalloc(g->first, (size_t)g->cells_cap + 1));
            GE_HIP_CHECK(c, dalloc(g->ord, cap));
            GE_HIP_CHECK(c, dalloc(g->frame, 8));
            GE_HIP_CHECK(c, dalloc(g->pk0, cap));
            GE_HIP_CHECK(c, dalloc(g->pk1, cap));
            GE_HIP_CHECK(c, dalloc(g->pkf, cap));
            GE_HIP_CHECK(c, dalloc(g->xy, cap));
        }
    }
    return GE_OK;
}

int bits_for(uint64_t v) {  // bits to hold values 0 .. v-1
    int b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

// System offsets, the per-slot system table and the reduction chunks.  Sets
// c->n; waits for the context's stream first (earlier frames read the tables).
int set_systems(ge_ctx *c, uint32_t nsys, const uint32_t *counts) {
    std::vector<uint32_t> off(nsys + 1, 0);
    uint32_t maxc = 1;
    for (uint32_t s = 0; s < nsys; ++s) {
        off[s + 1] = off[s] + counts[s];
        if (counts[s] > maxc) maxc = counts[s];
    }
    const uint32_t n = off[nsys];
    int idx_bits = kIdxBits, sys_bits = 0;
    if (nsys > 1) {
        idx_bits = bits_for(maxc);
        sys_bits = bits_for(nsys);
        if (idx_bits + sys_bits > kIdxBits) return GE_E_CAPACITY;
    }
    std::vector<uint32_t> chunk_lo, chunk_first(nsys + 1);
    for (uint32_t s = 0; s < nsys; ++s) {
        chunk_first[s] = (uint32_t)chunk_lo.size();
        for (uint32_t a = off[s]; a < off[s + 1]; a += kChunk) chunk_lo.push_back(a);
        if (counts[s] == 0) chunk_lo.push_back(off[s]);
    }
    chunk_first[nsys] = (uint32_t)chunk_lo.size();
    if (chunk_lo.size() > c->b.max_chunks) return GE_E_CAPACITY;
    // the walk's waves start at each system's first slot, so a wave holds the
    // same particles whether its system runs alone or in a batch
    std::vector<uint32_t> wave_off(nsys + 1, 0);
    for (uint32_t s = 0; s < nsys; ++s) wave_off[s + 1] = wave_off[s] + (counts[s] + 63u) / 64u;
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->idx_bits = idx_bits;
    c->sys_bits = sys_bits;
    c->nchunks = (uint32_t)chunk_lo.size();
    chunk_lo.push_back(n);
    c->nsys = nsys;
    c->n = n;
    c->host_off = off;
    GE_HIP_CHECK(c, hipMemcpy(c->b.sys_off, off.data(), sizeof(uint32_t) * (nsys + 1), hipMemcpyHostToDevice));
    GE_HIP_CHECK(c, hipMemcpy(c->b.chunk_lo, chunk_lo.data(), sizeof(uint32_t) * chunk_lo.size(), hipMemcpyHostToDevice));
    GE_HIP_CHECK(c, hipMemcpy(c->b.chunk_first, chunk_first.data(), sizeof(uint32_t) * (nsys + 1), hipMemcpyHostToDevice));
    GE_HIP_CHECK(c, hipMemcpy(c->b.sys_wave_off, wave_off.data(), sizeof(uint32_t) * (nsys + 1), hipMemcpyHostToDevice));
    c->nwaves = wave_off[nsys];
    if (n > 0) launch_sys_of(c);
    return GE_OK;
}

// Writes host arrays p[0 .. cnt) into slots [at, at + cnt) of state s.  NULL
// arrays give zero (sph_mass: 1; id: at-relative index 0 .. cnt-1).
int put_fields(ge_ctx *c, State &s, uint32_t at, uint32_t cnt, const ge_particles *p) {
    if (cnt == 0) return GE_OK;
    const size_t fb = sizeof(float) * cnt;
    auto put = [&](float *dst, const float *src, float fill) -> int {
        if (!dst) return GE_OK;
        if (src) {
            GE_HIP_CHECK(c, hipMemcpyAsync(dst + at, src, fb, hipMemcpyHostToDevice, c->stream));
        } else if (fill == 0.0f) {
            GE_HIP_CHECK(c, hipMemsetAsync(dst + at, 0, fb, c->stream));
        } else {
            std::vector<float> tmp(cnt, fill);
            GE_HIP_CHECK(c, hipMemcpyAsync(dst + at, tmp.data(), fb, hipMemcpyHostToDevice, c->stream));
            GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        }
        return GE_OK;
    };
    int r = 0;
    if ((r = put(s.x, p->x, 0.f)) || (r = put(s.y, p->y, 0.f)) || (r = put(s.vx, p->vx, 0.f)) ||
        (r = put(s.vy, p->vy, 0.f)) || (r = put(s.ax, p->ax, 0.f)) || (r = put(s.ay, p->ay, 0.f)) ||
        (r = put(s.pvx, p->pvx, 0.f)) || (r = put(s.pvy, p->pvy, 0.f)) || (r = put(s.m, p->mass, 0.f)))
        return r;
    if (p->flags) {
        GE_HIP_CHECK(c, hipMemcpyAsync(s.flags + at, p->flags, sizeof(uint32_t) * cnt, hipMemcpyHostToDevice, c->stream));
    } else {
        GE_HIP_CHECK(c, hipMemsetAsync(s.flags + at, 0, sizeof(uint32_t) * cnt, c->stream));
    }
    if (p->id) {
        GE_HIP_CHECK(c, hipMemcpyAsync(s.id + at, p->id, sizeof(uint32_t) * cnt, hipMemcpyHostToDevice, c->stream));
    } else {
        std::vector<uint32_t> ids(cnt);
        for (uint32_t i = 0; i < cnt; ++i) ids[i] = i;
        GE_HIP_CHECK(c, hipMemcpyAsync(s.id + at, ids.data(), sizeof(uint32_t) * cnt, hipMemcpyHostToDevice, c->stream));
        GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    }
    if (c->sph) {
        if ((r = put(s.sph_mass, p->sph_mass, 1.f)) || (r = put(s.rest_dens, p->rest_dens, 0.f)) ||
            (r = put(s.stiff, p->stiff, 0.f)) || (r = put(s.visc, p->visc, 0.f)) ||
            (r = put(s.cohesion, p->cohesion, 0.f)) || (r = put(s.dens, p->dens, 0.f)) ||
            (r = put(s.press, p->press, 0.f)) || (r = put(s.press_tmp, p->press_tmp, 0.f)) ||
            (r = put(s.pred_dens, p->pred_dens, 0.f)) || (r = put(s.ppx, p->ppx, 0.f)) ||
            (r = put(s.ppy, p->ppy, 0.f)) || (r = put(s.pfx, p->pfx, 0.f)) || (r = put(s.pfy, p->pfy, 0.f)))
            return r;
    }
    // ParticleRendering() defaults: white, size 1, no neighbours (particle.h:101-108)
    if (c->render) launch_render_fill(c->stream, s.color, s.pcolor, s.scolor, s.size, s.prev_size, s.nbrs, at, cnt,
                                      0xFFFFFFFFu, 1.0f);
    if (c->thermal) launch_thermal_fill(c, s, at, cnt);  // temp 288, no material
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

// Reads slots [at, at + cnt) of the current state into host arrays p[0 .. cnt).
int get_fields(ge_ctx *c, uint32_t at, uint32_t cnt, const ge_particles *p) {
    if (cnt == 0) return GE_OK;
    const State &s = c->b.st[c->b.cur];
    const size_t fb = sizeof(float) * cnt;
    auto get = [&](float *dst, const float *src) -> int {
        if (dst && src) GE_HIP_CHECK(c, hipMemcpyAsync(dst, src + at, fb, hipMemcpyDeviceToHost, c->stream));
        return GE_OK;
    };
    int r = 0;
    if ((r = get(p->x, s.x)) || (r = get(p->y, s.y)) || (r = get(p->vx, s.vx)) || (r = get(p->vy, s.vy)) ||
        (r = get(p->ax, s.ax)) || (r = get(p->ay, s.ay)) || (r = get(p->pvx, s.pvx)) || (r = get(p->pvy, s.pvy)) ||
        (r = get(p->mass, s.m)))
        return r;
    if (p->flags) GE_HIP_CHECK(c, hipMemcpyAsync(p->flags, s.flags + at, sizeof(uint32_t) * cnt, hipMemcpyDeviceToHost, c->stream));
    if (p->id) GE_HIP_CHECK(c, hipMemcpyAsync(p->id, s.id + at, sizeof(uint32_t) * cnt, hipMemcpyDeviceToHost, c->stream));
    if (c->sph) {
        if ((r = get(p->sph_mass, s.sph_mass)) || (r = get(p->rest_dens, s.rest_dens)) || (r = get(p->stiff, s.stiff)) ||
            (r = get(p->visc, s.visc)) || (r = get(p->cohesion, s.cohesion)) || (r = get(p->dens, s.dens)) ||
            (r = get(p->press, s.press)) || (r = get(p->press_tmp, s.press_tmp)) ||
            (r = get(p->pred_dens, s.pred_dens)) || (r = get(p->ppx, s.ppx)) || (r = get(p->ppy, s.ppy)) ||
            (r = get(p->pfx, s.pfx)) || (r = get(p->pfy, s.pfy)))
            return r;
    }
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int fail(ge_ctx *c, int code, const char *msg) {
    if (c) c->err = msg;
    return code;
}

int check_params(ge_ctx *c, const ge_step_params *p) {
    if (!p) return fail(c, GE_E_INVALID, "params is NULL");
    if (p->temperature && !c->thermal) return fail(c, GE_E_INVALID, "context created without enable_temperature");
    if (p->constraints && c->nsys > 1) return fail(c, GE_E_UNSUPPORTED, "constraints run on single-system contexts");
    if (p->temperature && p->accuracy)
        return fail(c, GE_E_UNSUPPORTED, "the heat transfer follows the reference walk (accuracy 0)");
    if ((p->periodic || p->sph_ground) && !(p->domain_w > 0.0f && p->domain_h > 0.0f))
        return fail(c, GE_E_INVALID, "periodic and ground modes need domain_w, domain_h > 0");
    if (p->sph && !c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    if (p->sph && c->nsys > 1) return fail(c, GE_E_UNSUPPORTED, "SPH runs on single-system contexts");
    if (p->sph && !(p->sph_cell > 0.0f && p->sph_h > 0.0f)) return fail(c, GE_E_INVALID, "sph_h and sph_cell must be > 0");
    if (p->accuracy < 0 || p->accuracy > 3) return fail(c, GE_E_INVALID, "accuracy must be 0, 1, 2 or 3");
    if (p->accuracy && !(p->theta > 0.0f)) return fail(c, GE_E_INVALID, "accuracy mode needs theta > 0");
    if (p->leaf_capacity > 64) return fail(c, GE_E_INVALID, "leaf_capacity must be 0 (reference rule) or 1..64");
    if (p->sph_search != 0 && p->sph_search != 1) return fail(c, GE_E_INVALID, "sph_search must be 0 or 1");
    if (p->sph && p->sph_search == 1 && !(p->sph_search_radius > 0.0f && p->sph_search_radius <= 8.0f * p->sph_cell))
        return fail(c, GE_E_INVALID, "sph_search_radius must be in (0, 8 cells]");
    if (p->walk_block != 0 && p->walk_block != 64 && p->walk_block != 128 && p->walk_block != 256)
        return fail(c, GE_E_INVALID, "walk_block must be 0, 64, 128 or 256");
    return GE_OK;
}

int check_device_errors(ge_ctx *c) {
    uint32_t e = 0;
    GE_HIP_CHECK(c, hipMemcpyAsync(&e, c->b.counters + kCntError, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
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

// Records a hipEvent pair around one phase when profiling is on.
struct PhaseScope {
    ge_ctx *c;
    ge_ctx::PhaseEvent ev{};
    bool on;
    PhaseScope(ge_ctx *ctx, int phase) : c(ctx), on(ctx->prof) {
        if (!on) return;
        ev.phase = phase;
        if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) {
            on = false;
            return;
        }
        (void)hipEventRecord(ev.a, c->stream);
    }
    ~PhaseScope() {
        if (!on) return;
        (void)hipEventRecord(ev.b, c->stream);
        c->prof_events.push_back(ev);
    }
};

bool use_direct(const ge_ctx *c, const ge_step_params *p) {
    // globalLogic.cpp:1319-1348: naive flag, or N < 5000 on either code path
    if (c->nsys > 1) return false;  // the batched form always walks each system's tree
    return p->naive || (int64_t)c->n < (int64_t)p->direct_below;
}

// createConstraints runs each frame of a context that keeps temperature
// fields (hasSolidified is set by temperatureCalculation) and constraints;
// it reads the created count back, so such frames are not captured in graphs.
bool creates_constraints(const ge_ctx *c, const ge_step_params *p) {
    return p->constraints && c->thermal && c->cons_cap > 0 && c->b.ccreate;
}

// One frame of updateScene() (globalLogic.cpp:1155-1372).
int frame(ge_ctx *c, const ge_step_params *p, float dt_start, bool count_inter) {
    c->leaf_cap = p->leaf_capacity;
    if (dt_start != 0.0f) {
        // integrateStart + boundingBox, then computeMortonKeys, the sort and the tree
        {
            PhaseScope s(c, GE_PHASE_INTEGRATE_BBOX);
            launch_integrate_start_bbox(c, dt_start, p->sph, p->sph_max_vel, 1, domain_of(p));
        }
        { PhaseScope s(c, GE_PHASE_KEYS); launch_morton_keys(c); }
        { PhaseScope s(c, GE_PHASE_SORT); launch_sort_keys(c); }
        { PhaseScope s(c, GE_PHASE_GATHER); launch_gather(c); }
        if (!p->naive) {
            PhaseScope s(c, GE_PHASE_TREE);
            launch_tree_build(c, true);
            c->have_tree = true;
        }
    }
    bool grid_done = false;
    if (creates_constraints(c, p)) {
        // createConstraints (globalLogic.cpp:1299-1300) queries the grid the
        // frame built on the sorted positions, before the forces
        PhaseScope s(c, GE_PHASE_SPH_GRID);
        if (dt_start == 0.0f) launch_integrate_start_bbox(c, 0.0f, 0, 0.0f, 0);  // the grid frame's extremes
        if (int r = sph_new_grid(c, p)) return r;
        grid_done = true;
        uint32_t created = 0;
        if (int r = sph_create_constraints(c, p, &created)) return r;
        if (created) {
            if (int r = cons_prepare(c)) return r;
        }
    }
    if (p->dt == 0.0f) return GE_OK;
    if (p->sph && !grid_done) {
        PhaseScope s(c, GE_PHASE_SPH_GRID);
        if (int r = sph_new_grid(c, p)) return r;
    }
    const bool temp = p->temperature != 0;  // temperatureCalculation runs before integrateEnd
    const bool cons = p->constraints && c->cons_bound > 0;  // Physics::constraints after PCISPH
    const bool late = temp || cons;  // work between the forces and integrateEnd
    const float end_kick = (p->sph || late) ? 0.0f : p->dt * 0.5f;  // fused integrateEnd when nothing follows
    const bool ground = p->sph && p->sph_ground;           // groundModeBoundary runs before integrateEnd
    // the walk runs for gravity or for the heat transfer (globalLogic.cpp:1316)
    if (p->G != 0.0f || temp) {
        if (use_direct(c, p) || !c->have_tree) {
            PhaseScope s(c, GE_PHASE_GRAVITY);
            launch_direct(c, p, end_kick);  // naiveGravityAVX2 moves no heat
        } else {
            if (p->accuracy) {
                PhaseScope s(c, GE_PHASE_MULTIPOLE);
                launch_multipoles(c, p->theta);
            }
            if (temp) launch_node_temps(c);
            PhaseScope s(c, GE_PHASE_GRAVITY);
            launch_bh(c, p, end_kick, count_inter);
        }
    } else {
        launch_zero_acc(c);
        if (!p->sph && !late) {
            PhaseScope s(c, GE_PHASE_INTEGRATE_END);
            launch_integrate_end(c, p->dt);
        }
    }
    if (p->sph) {
        PhaseScope s(c, GE_PHASE_SPH);
        if (int r = sph_pcisph(c, p, (ground || late) ? 0.0f : p->dt)) return r;
        if (ground) launch_sph_ground(c, p);
    }
    if (cons) {
        if (int r = cons_solve(c, p, false)) return r;
    }
    if (temp) launch_temperature(c, p);
    if (ground || late) {
        PhaseScope s(c, GE_PHASE_INTEGRATE_END);
        launch_integrate_end(c, p->dt);
    }
    return GE_OK;
}

void drop_graph(ge_ctx *c) {
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    c->graph = ge_ctx::FrameGraph{};
}

bool graph_matches(const ge_ctx *c, const ge_step_params *p) {
    const ge_ctx::FrameGraph &g = c->graph;
    return g.exec && g.layout_version == c->layout_version && g.cur == c->b.cur && g.sort_a == c->b.sort_a &&
           g.grid_a == c->b.grid_pos.keys && g.grid_b == c->b.grid_pred.keys &&
           std::memcmp(&g.params, p, sizeof *p) == 0;
}

// Captures two frames at dt (the pair leaves every alternating buffer where it
// found it, so the graph can be replayed back to back).  Capture runs no work.
int capture_pair(ge_ctx *c, const ge_step_params *p) {
    drop_graph(c);
    ge_ctx::FrameGraph g;
    g.params = *p;
    g.layout_version = c->layout_version;
    g.cur = c->b.cur;
    g.sort_a = c->b.sort_a;
    g.grid_a = c->b.grid_pos.keys;
    g.grid_b = c->b.grid_pred.keys;
    GE_HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = frame(c, p, p->dt, false);
    if (r == GE_OK) r = frame(c, p, p->dt, false);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (r != GE_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
    }
    if (e != hipSuccess || !graph) return fail(c, GE_E_HIP, "hipStreamEndCapture failed");
    const bool same = c->b.cur == g.cur && c->b.sort_a == g.sort_a && c->b.grid_pos.keys == g.grid_a &&
                      c->b.grid_pred.keys == g.grid_b;
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(c, GE_E_HIP, "hipGraphInstantiate failed");
    if (!same) {  // buffers did not return to their starting places: no replay
        (void)hipGraphExecDestroy(exec);
        return fail(c, GE_E_HIP, "frame pair does not restore the buffer layout");
    }
    g.exec = exec;
    c->graph = g;
    return GE_OK;
}

}  // namespace

extern "C" {

const char *ge_version(void) { return "galaxy-step-mi355x 0.1 (gfx950)"; }

void ge_default_config(ge_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->capacity = 1u << 20;
    cfg->device = 0;
    cfg->enable_sph = 0;
    cfg->max_systems = 1;
}

void ge_default_params(ge_step_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->G = 6.674e-11f;       // parameters.h:114 (times gravityMultiplier 1)
    p->theta = 0.8f;         // parameters.h:117
    p->softening = 1.5f;     // parameters.h:116
    p->dt_start = 0.045f;    // fixedDeltaTime (parameters.h:129)
    p->dt = 0.045f;
    p->naive = 0;
    p->direct_below = 5000;  // globalLogic.cpp:1323
    p->sph = 0;
    p->sph_h = 3.0f;         // SPH.h:17
    p->sph_cell = 3.0f;      // neighborSearch.h:604
    p->sph_unit_mass = 0.03f;  // parameters.h:311
    p->viscosity = 0.3f;
    p->cohesion_coef = 1.0f;
    p->delta = 19000.0f;
    p->stiff_mult = 1.0f;
    p->sph_max_vel = 250.0f;   // parameters.h:118
    p->sph_search = 0;
    p->sph_search_radius = 3.5f;  // NeighborSearchV2::searchRadius (neighborSearch.h:441)
    p->domain_w = 3840.0f;        // domainSize (parameters.h:97)
    p->domain_h = 2160.0f;
    p->vertical_gravity = 3.0f;   // parameters.h:316
    p->boundary_friction = 0.0f;  // parameters.h:361
    p->compensated = 1;
}

int ge_create(ge_ctx **out, const ge_config *cfg) {
    if (!out || !cfg) return GE_E_INVALID;
    *out = nullptr;
    if (cfg->capacity == 0 || cfg->capacity > kMaxParticles) return GE_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GE_E_NODEVICE;
    if (cfg->device < 0 || cfg->device >= ndev) return GE_E_NODEVICE;
    ge_ctx *c = new (std::nothrow) ge_ctx();
    if (!c) return GE_E_NOMEM;
    c->device = cfg->device;
    c->cap = cfg->capacity;
    c->sph = cfg->enable_sph != 0;
    c->render = cfg->enable_render != 0;
    c->thermal = cfg->enable_temperature != 0;
    c->precise = cfg->precise_positions != 0;
    c->cons_cap = cfg->max_constraints;
    ge_default_constraint_params(&c->cons_p);
    if (c->thermal && !c->sph) {  // the materials write the SPH fields
        delete c;
        return GE_E_INVALID;
    }
    c->max_sys = cfg->max_systems ? cfg->max_systems : 1;
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GE_E_HIP;
    }
    // the fixed scratch of the entry point that allocates nothing, once per device
    if (int r = noalloc_setup(c->device)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return r;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
        c->cus = (uint32_t)cus;
    t_alloc_bytes = 0;
    int ra = alloc_all(c);
    c->device_bytes = t_alloc_bytes;
    if (ra == GE_OK && c->thermal) {  // the reference's constants and materials until ge_set_thermal
        ge_thermal_params tp;
        ge_default_thermal_params(&tp);
        ge_material mats[GE_MAX_MATERIALS];
        uint32_t nm = 0;
        ge_default_materials(mats, GE_MAX_MATERIALS, &nm);
        ra = ge_set_thermal(c, &tp, mats, nm);
    }
    if (int r = ra) {
        free_all(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return r == GE_E_HIP ? GE_E_NOMEM : r;
    }
    *out = c;
    return GE_OK;
}

int ge_destroy(ge_ctx *c) {
    if (!c) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->prof_events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    drop_graph(c);
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GE_OK;
}

const char *ge_last_error(const ge_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ge_memory_bytes(const ge_ctx *c, uint64_t *bytes) {
    if (!c || !bytes) return GE_E_INVALID;
    *bytes = c->device_bytes;
    return GE_OK;
}

int ge_count(const ge_ctx *c, uint32_t *n) {
    if (!c || !n) return GE_E_INVALID;
    *n = c->n;
    return GE_OK;
}

int ge_sync(ge_ctx *c) {
    if (!c) return GE_E_INVALID;
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return check_device_errors(c);
}

int ge_upload(ge_ctx *c, uint32_t n, const ge_particles *p) { return ge_upload_systems(c, 1, &n, p); }

int ge_upload_systems(ge_ctx *c, uint32_t nsys, const uint32_t *counts, const ge_particles *p) {
    if (!c || !p || !counts || nsys == 0) return GE_E_INVALID;
    if (nsys > c->max_sys) return fail(c, GE_E_CAPACITY, "more systems than max_systems");
    uint64_t total = 0;
    for (uint32_t s = 0; s < nsys; ++s) {
        if (nsys > 1 && counts[s] == 0) return fail(c, GE_E_INVALID, "every system needs a particle");
        total += counts[s];
    }
    if (total > c->cap) return fail(c, GE_E_CAPACITY, "particle count above capacity");
    if (total > 0 && (!p->x || !p->y || !p->mass)) return fail(c, GE_E_INVALID, "x, y and mass are required");
    (void)hipSetDevice(c->device);
    if (int r = set_systems(c, nsys, counts)) {
        return fail(c, r, r == GE_E_CAPACITY ? "systems too many or too large for one packed sort word" : "HIP error");
    }
    c->b.cur = 0;
    c->layout_version += 1;
    c->have_tree = false;
    c->have_grid = false;
    return put_fields(c, c->b.st[0], 0, c->n, p);
}

int ge_systems_update(ge_ctx *c, uint32_t nkeep, const uint32_t *keep, uint32_t nadd, const uint32_t *counts,
                      const ge_particles *p) {
    if (!c || (nkeep && !keep) || (nadd && (!counts || !p))) return GE_E_INVALID;
    const uint32_t old_nsys = c->n ? c->nsys : 0;
    const std::vector<uint32_t> old_off = c->host_off;
    for (uint32_t t = 0; t < nkeep; ++t) {
        if (keep[t] >= old_nsys || (t > 0 && keep[t] <= keep[t - 1]))
            return fail(c, GE_E_INVALID, "keep[] must list existing systems in ascending order");
    }
    const uint32_t nsys = nkeep + nadd;
    if (nsys == 0) {
        c->n = 0;
        c->nsys = 1;
        c->host_off.assign(2, 0);
        c->layout_version += 1;
        c->have_tree = c->have_grid = false;
        return GE_OK;
    }
    if (nsys > c->max_sys) return fail(c, GE_E_CAPACITY, "more systems than max_systems");
    if (c->sph && nsys > 1) return fail(c, GE_E_UNSUPPORTED, "SPH contexts hold one system");
    std::vector<uint32_t> cnt(nsys), src(nkeep + 1, 0);
    uint64_t total = 0;
    bool moved = false;
    uint32_t n_kept = 0;
    for (uint32_t t = 0; t < nkeep; ++t) {
        cnt[t] = old_off[keep[t] + 1] - old_off[keep[t]];
        src[t] = old_off[keep[t]];
        if (src[t] != n_kept) moved = true;
        n_kept += cnt[t];
    }
    for (uint32_t a = 0; a < nadd; ++a) {
        if (counts[a] == 0) return fail(c, GE_E_INVALID, "every system needs a particle");
        cnt[nkeep + a] = counts[a];
    }
    for (uint32_t t = 0; t < nsys; ++t) total += cnt[t];
    if (total > c->cap) return fail(c, GE_E_CAPACITY, "particle count above capacity");
    if (nadd && total > n_kept && (!p->x || !p->y || !p->mass)) return fail(c, GE_E_INVALID, "x, y and mass are required");
    (void)hipSetDevice(c->device);
    // the new layout must fit a packed sort word before anything moves
    {
        uint32_t maxc = 1;
        for (uint32_t v : cnt) maxc = v > maxc ? v : maxc;
        if (nsys > 1 && bits_for(maxc) + bits_for(nsys) > kIdxBits)
            return fail(c, GE_E_CAPACITY, "systems too many or too large for one packed sort word");
    }
    if (int r = set_systems(c, nsys, cnt.data())) return fail(c, r, "HIP error");
    if (moved) {
        GE_HIP_CHECK(c, hipMemcpy(c->b.sys_src, src.data(), sizeof(uint32_t) * nkeep, hipMemcpyHostToDevice));
        launch_compact(c, n_kept);
    }
    c->layout_version += 1;
    c->have_tree = false;
    c->have_grid = false;
    if (nadd) return put_fields(c, c->b.st[c->b.cur], n_kept, (uint32_t)(total - n_kept), p);
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge_download(ge_ctx *c, const ge_particles *p) {
    if (!c || !p) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    if (int r = get_fields(c, 0, c->n, p)) return r;
    return check_device_errors(c);
}

int ge_download_system(ge_ctx *c, uint32_t sys, const ge_particles *p) {
    if (!c || !p) return GE_E_INVALID;
    if (c->n == 0 || sys >= c->nsys) return fail(c, GE_E_INVALID, "no such system");
    (void)hipSetDevice(c->device);
    if (int r = get_fields(c, c->host_off[sys], c->host_off[sys + 1] - c->host_off[sys], p)) return r;
    return check_device_errors(c);
}

int ge_poll(ge_ctx *c, int32_t *done) {
    if (!c || !done) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) {
        *done = 1;
        return GE_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return GE_OK;
    }
    c->err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
    return GE_E_HIP;
}

int ge_step(ge_ctx *c, const ge_step_params *p, int32_t nsteps) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (nsteps < 0) return fail(c, GE_E_INVALID, "nsteps < 0");
    if (c->n == 0 || nsteps == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    if (p->accuracy) {
        if (int r = ensure_multipole_buffers(c)) return r;
    }
    if (p->constraints) {  // colouring (host round trips) before any frame or capture
        if (int r = cons_prepare(c)) return r;
    }
    int32_t k = 0;
    if (p->use_graph && !c->prof && nsteps >= 3 && !creates_constraints(c, p)) {
        // the first frame integrates with dt_start, so it runs eagerly
        if (int r = frame(c, p, p->dt_start, false)) return r;
        k = 1;
        const int32_t pairs = (nsteps - 1) / 2;
        if (!graph_matches(c, p)) {
            if (int r = capture_pair(c, p)) return r;
        }
        for (int32_t q = 0; q < pairs; ++q) GE_HIP_CHECK(c, hipGraphLaunch(c->graph.exec, c->stream));
        k += 2 * pairs;
    }
    for (; k < nsteps; ++k) {
        const float dt0 = (k == 0) ? p->dt_start : p->dt;
        if (int r = frame(c, p, dt0, false)) return r;
    }
    GE_HIP_CHECK(c, hipGetLastError());
    return GE_OK;
}

int ge_system_offsets(const ge_ctx *c, uint32_t *nsys, uint32_t *offsets, uint32_t cap) {
    if (!c || !nsys) return GE_E_INVALID;
    *nsys = c->nsys;
    if (offsets) {
        if (cap < c->nsys + 1) return GE_E_CAPACITY;
        for (uint32_t s = 0; s <= c->nsys; ++s) offsets[s] = c->host_off.empty() ? 0 : c->host_off[s];
    }
    return GE_OK;
}

int ge_bounding_box(ge_ctx *c, float out_bb[3]) {
    if (!c || !out_bb) return GE_E_INVALID;
    if (c->n == 0) return fail(c, GE_E_INVALID, "no particles");
    if (c->nsys != 1) return fail(c, GE_E_INVALID, "stage calls take a single-system context");
    (void)hipSetDevice(c->device);
    launch_integrate_start_bbox(c, 0.0f, 0, 0.0f, 0);
    float h[4];
    GE_HIP_CHECK(c, hipMemcpyAsync(h, c->b.bb, sizeof h, hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    out_bb[0] = h[0]; out_bb[1] = h[1]; out_bb[2] = h[2];
    return GE_OK;
}

int ge_compute_morton_keys(ge_ctx *c, const float bb[3]) {
    if (!c || !bb) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    if (c->nsys != 1) return fail(c, GE_E_INVALID, "stage calls take a single-system context");
    (void)hipSetDevice(c->device);
    float h[4] = {bb[0], bb[1], bb[2], 0.0f};
    GE_HIP_CHECK(c, hipMemcpyAsync(c->b.bb, h, sizeof h, hipMemcpyHostToDevice, c->stream));
    launch_morton_keys(c);
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge_sort_by_morton_key(ge_ctx *c) {
    if (!c) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch_sort_keys(c);
    launch_gather(c);
    c->have_tree = false;
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge_build_quadtree(ge_ctx *c, const float bb[3]) {
    if (!c || !bb) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    if (c->nsys != 1) return fail(c, GE_E_INVALID, "stage calls take a single-system context");
    (void)hipSetDevice(c->device);
    float h[4] = {bb[0], bb[1], bb[2], 0.0f};
    GE_HIP_CHECK(c, hipMemcpyAsync(c->b.bb, h, sizeof h, hipMemcpyHostToDevice, c->stream));
    c->leaf_cap = 0;  // the reference's leaf rule
    launch_tree_build(c);
    c->have_tree = true;
    return ge_sync(c);
}

int ge_compute_gravity(ge_ctx *c, const ge_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    if (p->G == 0.0f) {
        launch_zero_acc(c);
    } else if (use_direct(c, p) || !c->have_tree) {
        launch_direct(c, p, 0.0f);
    } else {
        if (p->accuracy) {
            if (int r = ensure_multipole_buffers(c)) return r;
            launch_multipoles(c, p->theta);
        }
        launch_bh(c, p, 0.0f, true);
    }
    return ge_sync(c);
}

int ge_integrate_start(ge_ctx *c, const ge_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch_integrate_start_bbox(c, p->dt_start, p->sph, p->sph_max_vel, 1, domain_of(p));
    return ge_sync(c);
}

int ge_integrate_end(ge_ctx *c, const ge_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch_integrate_end(c, p->dt);
    return ge_sync(c);
}

int ge_sph_new_grid(ge_ctx *c, const ge_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (!c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    // the grid frame uses the current bounding box
    launch_integrate_start_bbox(c, 0.0f, 0, 0.0f, 0);
    if (int r = sph_new_grid(c, p)) return r;
    return ge_sync(c);
}

int ge_sph_pcisph(ge_ctx *c, const ge_step_params *p) {
    if (!c) return GE_E_INVALID;
    if (int r = check_params(c, p)) return r;
    if (!c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    if (!c->have_grid) return fail(c, GE_E_INVALID, "ge_sph_new_grid must run first");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    if (int r = sph_pcisph(c, p, 0.0f)) return r;
    if (p->sph_ground) launch_sph_ground(c, p);  // pcisphSolver (SPH.h:623-632)
    return ge_sync(c);
}

int ge_get_keys(ge_ctx *c, uint64_t *keys) {
    if (!c || !keys) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    GE_HIP_CHECK(c, hipMemcpyAsync(keys, c->b.keys, sizeof(uint64_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge_get_perm(ge_ctx *c, uint32_t *perm) {
    if (!c || !perm) return GE_E_INVALID;
    if (c->n == 0) return GE_OK;
    GE_HIP_CHECK(c, hipMemcpyAsync(perm, c->b.perm, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge_get_node_count(ge_ctx *c, uint32_t *count) {
    if (!c || !count) return GE_E_INVALID;
    *count = 0;
    if (!c->have_tree || c->n == 0) return GE_OK;
    GE_HIP_CHECK(c, hipMemcpyAsync(count, c->b.counters + kCntNodeTotal, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return check_device_errors(c);
}

int ge_get_nodes(ge_ctx *c, ge_node *out, uint32_t cap) {
    if (!c || !out) return GE_E_INVALID;
    uint32_t nn = 0;
    if (int r = ge_get_node_count(c, &nn)) return r;
    if (nn > cap) return fail(c, GE_E_CAPACITY, "node buffer too small");
    if (!c->debug) return fail(c, GE_E_INVALID, "node geometry not kept");
    std::vector<NodeRec> rec(nn);
    std::vector<float4> geo(nn);
    if (nn) {
        GE_HIP_CHECK(c, hipMemcpy(rec.data(), c->b.nodes, sizeof(NodeRec) * nn, hipMemcpyDeviceToHost));
        GE_HIP_CHECK(c, hipMemcpy(geo.data(), c->b.node_geom, sizeof(float4) * nn, hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < nn; ++i) {
        ge_node &o = out[i];
        o.posx = geo[i].x; o.posy = geo[i].y; o.size = geo[i].z;
        o.comx = rec[i].comx; o.comy = rec[i].comy; o.mass = rec[i].mass;
        o.start = rec[i].start; o.end = rec[i].end; o.next = rec[i].next;
        uint32_t ch = i + 1;
        for (int q = 0; q < 4; ++q) {
            if (rec[i].child_mask & (1u << q)) {
                o.child[q] = ch;
                ch += (ch < nn ? rec[ch].next : 0u) + 1;
            } else {
                o.child[q] = kNone;
            }
        }
    }
    return GE_OK;
}

int ge_get_interactions(ge_ctx *c, uint64_t *count) {
    if (!c || !count) return GE_E_INVALID;
    GE_HIP_CHECK(c, hipMemcpyAsync(count, c->b.interactions, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge_get_walk_stats(ge_ctx *c, uint64_t out[4]) {
    if (!c || !out) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    GE_HIP_CHECK(c, hipMemcpyAsync(out, c->b.interactions, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

int ge_get_interaction_counts(ge_ctx *c, uint64_t *nodes, uint64_t *pairs) {
    if (!c || !nodes || !pairs) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    uint64_t h[2] = {0, 0};
    GE_HIP_CHECK(c, hipMemcpyAsync(h, c->b.interactions, sizeof h, hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    *nodes = h[1];
    *pairs = h[0] - h[1];
    return GE_OK;
}

int ge_profile_enable(ge_ctx *c, int enable) {
    if (!c) return GE_E_INVALID;
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (auto &e : c->prof_events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->prof_events.clear();
    c->prof = enable != 0;
    return GE_OK;
}

int ge_profile_read(ge_ctx *c, double ms[GE_NPHASES], uint32_t launches[GE_NPHASES]) {
    if (!c || !ms || !launches) return GE_E_INVALID;
    for (int k = 0; k < GE_NPHASES; ++k) { ms[k] = 0.0; launches[k] = 0; }
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (auto &e : c->prof_events) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, e.a, e.b) == hipSuccess && e.phase >= 0 && e.phase < GE_NPHASES) {
            ms[e.phase] += t;
            launches[e.phase] += 1;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->prof_events.clear();
    return GE_OK;
}

int ge_neighbor_counts(ge_ctx *c, const ge_step_params *p, uint32_t *counts) {
    if (!c || !p || !counts) return GE_E_INVALID;
    if (!c->sph) return fail(c, GE_E_INVALID, "context created without enable_sph");
    if (!(p->sph_cell > 0.0f && p->sph_search_radius > 0.0f && p->sph_search_radius <= 8.0f * p->sph_cell))
        return fail(c, GE_E_INVALID, "sph_cell and sph_search_radius must be > 0 (radius at most 8 cells)");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch_integrate_start_bbox(c, 0.0f, 0, 0.0f, 0);  // the grid frame uses the current extremes
    return sph_neighbor_counts(c, p, counts);
}

int ge_spawn_galaxy(ge_ctx *c, const ge_galaxy_params *g, uint64_t seed, const float *uniforms, uint32_t *first) {
    if (!c || !g) return GE_E_INVALID;
    if (c->n && c->nsys > 1) return fail(c, GE_E_UNSUPPORTED, "the spawner appends to single-system contexts");
    if (!(g->scale_length > 0.0f && g->radius_core_dm > 0.0f && g->outer_radius_dm >= 0.0f &&
          std::isfinite(g->outer_radius) && std::isfinite(g->center_x) && std::isfinite(g->center_y)))
        return fail(c, GE_E_INVALID, "scale_length and radius_core_dm must be > 0, radii and centre finite");
    const uint64_t add = (uint64_t)g->visible + g->dark;
    if ((uint64_t)c->n + add > c->cap) return fail(c, GE_E_CAPACITY, "particle count above capacity");
    if (uniforms) {
        for (uint64_t k = 0; k < 3ull * g->visible + 5ull * g->dark; ++k)
            if (!(uniforms[k] >= 0.0f && uniforms[k] <= 1.0f)) return fail(c, GE_E_INVALID, "uniforms must lie in [0, 1]");
    }
    const uint32_t at = c->n;
    if (first) *first = at;
    if (add == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    float *du = nullptr;
    if (uniforms) {
        const size_t bytes = sizeof(float) * (3ull * g->visible + 5ull * g->dark);
        GE_HIP_CHECK(c, hipMalloc(&du, bytes));
        if (hipMemcpy(du, uniforms, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(du);
            return fail(c, GE_E_HIP, "uniform copy failed");
        }
    }
    const uint32_t total = (uint32_t)(at + add);
    if (int r = set_systems(c, 1, &total)) {
        if (du) (void)hipFree(du);
        return fail(c, r, "HIP error");
    }
    launch_spawn_galaxy(c, c->b.st[c->b.cur], at, g, seed, du);
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (du) (void)hipFree(du);
    if (e != hipSuccess) return fail(c, GE_E_HIP, "galaxy spawn failed");
    c->layout_version += 1;
    c->have_tree = false;
    c->have_grid = false;
    return GE_OK;
}

void ge_default_merger_params(ge_merger_params *m) {
    if (!m) return;
    std::memset(m, 0, sizeof *m);
    m->density_radius = 10.0f;     // globalLogic.cpp:1253-1254 once the merger runs
    m->cell_size = 3.0f;           // neighborSearch.h:11
    m->texture_half_size = 16.0f;  // parameters.h:161
    m->size_multiplier = 0.6f;     // parameters.h:202
}

int ge_merger_cell_size(ge_ctx *c, const ge_merger_params *m, float *cell) {
    if (!c || !m || !cell) return GE_E_INVALID;
    if (!c->render) return fail(c, GE_E_INVALID, "the merger reads the render fields: create with enable_render");
    (void)hipSetDevice(c->device);
    return merger_cell_size(c, m, cell);
}

int ge_merge(ge_ctx *c, const ge_merger_params *m, uint32_t *removed) {
    if (!c || !m) return GE_E_INVALID;
    if (removed) *removed = 0;
    if (!c->render) return fail(c, GE_E_INVALID, "the merger reads the render fields: create with enable_render");
    if (c->nsys > 1) return fail(c, GE_E_UNSUPPORTED, "the merger runs on single-system contexts");
    if (!(m->cell_size > 0.0f && m->density_radius >= 0.0f))
        return fail(c, GE_E_INVALID, "cell_size must be > 0 and density_radius >= 0");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    const uint32_t before = c->n;
    std::vector<uint32_t> dels;
    if (int r = merger_run(c, m, dels)) return r;
    if (dels.empty()) return GE_OK;
    uint32_t new_n = 0;
    if (int r = merger_apply(c, dels, new_n)) return r;
    if (int r = set_systems(c, 1, &new_n)) return fail(c, r, "HIP error");
    c->layout_version += 1;
    c->have_tree = false;
    c->have_grid = false;
    if (removed) *removed = before - new_n;
    return GE_OK;
}

int ge_get_neighbors(ge_ctx *c, int32_t which, int64_t *off, uint32_t *idx, int64_t cap, int64_t *total) {
    if (!c || !off || !total || (which != 0 && which != 1)) return GE_E_INVALID;
    if (!c->sph || !c->have_grid) return fail(c, GE_E_INVALID, "no SPH grid");
    (void)hipSetDevice(c->device);
    return sph_neighbors(c, which, off, idx, cap, total);
}

}  // extern "C"
