// ge_thermal.hip -- the temperature coupling of the 2D frame (include/ge_thermal.h):
// Node::gridTemp of the final tree (GE/include/Physics/quadtree.h:53-100),
// Physics::temperatureCalculation (GE/src/Physics/physics.cpp:567-652) and
// the temperature fields' defaults and transfers.  The heat transfer inside
// the walk is the TEMP form of k_bh (ge_step.hip).
//
// gridTemp bottom-up without levels: every leaf sums its particles' temp in
// order, then climbs its parent links; at each parent an arrival counter
// elects the last child to arrive, which sums the children's values in the
// order q = 0, 2, 1, 3 of computeInternalMass and climbs on.  The values are
// the reference's float sums whichever thread computes them.  Writers fence
// before their arrival; readers fence after winning it and load through the
// device-coherent path.  No thread waits on another, so the pass cannot
// hang.
#include <cstring>

#include "ge_internal.hpp"

namespace ge {
namespace {

__global__ __launch_bounds__(kBlock) void k_thermal_fill(float *__restrict__ temp, uint32_t *__restrict__ label,
                                                         float *__restrict__ ke, float *__restrict__ pke,
                                                         uint32_t at, uint32_t cnt) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cnt) return;
    const uint32_t i = at + k;
    temp[i] = 288.0f;  // ParticlePhysics(pos, vel, ...) (particle.h:61)
    label[i] = 0u;
    ke[i] = 0.0f;
    pke[i] = 0.0f;
}

__device__ __forceinline__ uint32_t node_total(const uint32_t *counters) {
    return (counters[kCntError] & (kErrNodeCap | kErrDepth)) ? 0u : counters[kCntNodeTotal];
}

__global__ __launch_bounds__(kBlock) void k_node_links_init(uint32_t *__restrict__ parent,
                                                            uint32_t *__restrict__ arrive, uint32_t cap) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cap) return;
    parent[k] = kNone;
    arrive[k] = 0u;
}

// parent[c] for the children of every internal node; children follow their
// parent in pre-order, in increasing quadrant q
__global__ __launch_bounds__(kBlock) void k_node_links(const NodeRec *__restrict__ nodes,
                                                       const uint32_t *__restrict__ counters,
                                                       uint32_t *__restrict__ parent) {
    const uint32_t nn = node_total(counters);
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= nn) return;
    const uint32_t mask = nodes[idx].child_mask;
    uint32_t c = idx + 1;
    for (uint32_t k = 0; k < (uint32_t)__popc(mask) && c < nn; ++k) {
        parent[c] = idx;
        c = c + nodes[c].next + 1;
    }
}

__device__ __forceinline__ float load_coherent(const float *p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t *>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void store_coherent(float *p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t *>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBlock) void k_node_temps_up(const NodeRec *__restrict__ nodes,
                                                          const uint32_t *__restrict__ counters,
                                                          const uint32_t *__restrict__ parent,
                                                          uint32_t *__restrict__ arrive, const float *__restrict__ temp,
                                                          uint32_t n, float *__restrict__ gtemp) {
#pragma clang fp contract(off)
    const uint32_t nn = node_total(counters);
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= nn) return;
    const NodeRec nd = nodes[idx];
    if (nd.child_mask != 0u) return;  // leaves start the climb
    // computeLeafMass's gridTemp (quadtree.h:53-67)
    float t = 0.0f;
    const uint32_t e = min(nd.end, n);
    for (uint32_t j = nd.start; j < e; ++j) t = t + temp[j];
    store_coherent(&gtemp[idx], t);
    uint32_t node = idx;
    for (;;) {
        const uint32_t par = parent[node];
        if (par == kNone || par >= nn) break;
        const uint32_t mask = nodes[par].child_mask;
        __threadfence();
        const uint32_t prev = atomicAdd(&arrive[par], 1u);
        if (prev + 1u != (uint32_t)__popc(mask)) break;  // a sibling is still on its way
        __threadfence();
        // computeInternalMass's gridTemp: children q = 0, 2, 1, 3
        uint32_t ch[4] = {kNone, kNone, kNone, kNone};
        uint32_t c = par + 1;
        for (int q = 0; q < 4; ++q) {
            if (!((mask >> q) & 1u) || c >= nn) continue;
            ch[q] = c;
            c = c + nodes[c].next + 1;
        }
        const int order[4] = {0, 2, 1, 3};
        float s = 0.0f;
        for (int k = 0; k < 4; ++k) {
            const uint32_t cq = ch[order[k]];
            if (cq != kNone) s = s + load_coherent(&gtemp[cq]);
        }
        store_coherent(&gtemp[par], s);
        node = par;
    }
}

__device__ __forceinline__ const ge_material *find_material(const ge_material *mats, uint32_t nmat, uint32_t id) {
    for (uint32_t k = 0; k < nmat; ++k)
        if (mats[k].id == id) return &mats[k];
    return nullptr;
}

// Physics::temperatureCalculation (physics.cpp:567-652)
__global__ __launch_bounds__(kBlock) void k_temperature(State s, uint32_t n, const ge_material *__restrict__ mats,
                                                        uint32_t nmat, float ambient, float rate, float dt,
                                                        float base_mass) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ge_material *mt = find_material(mats, nmat, s.label[i]);
    const float K = mt ? mt->heat_conductivity : 0.05f;
    const float vx = s.vx[i], vy = s.vy[i], pvx = s.pvx[i], pvy = s.pvy[i];
    const float v = sqrtf(vx * vx + vy * vy);
    const float pv = sqrtf(pvx * pvx + pvy * pvy);
    const float sm = s.sph_mass[i];
    const float ke = 0.5f * sm * v * v;
    const float pke = 0.5f * sm * pv * pv;
    s.ke[i] = ke;
    s.prev_ke[i] = pke;
    const float q = fabsf(ke - pke);
    const float dT = q / (2.0f * K * sm + 1.0f);
    float t = s.temp[i] + dT;
    const float diff = t - ambient;
    const float cool = -(K * rate) * diff * dt;
    t = t + cool;
    s.temp[i] = t;
    if (!mt) return;
    uint32_t f = s.flags[i];
    bool hot = (f & GE_FLAG_HOT_POINT) != 0, solid = (f & GE_FLAG_SOLIDIFIED) != 0;
    float nsm, rd, st, vi, co;
    if (t >= mt->hot_point) {
        nsm = mt->hot_mass_mult; rd = mt->hot_rest_dens; st = mt->hot_stiff; vi = mt->hot_visc; co = mt->hot_cohesion;
        if (mt->cold_point == 0.0f) hot = true;
    } else if (t <= mt->cold_point) {
        nsm = mt->cold_mass_mult; rd = mt->cold_rest_dens; st = mt->cold_stiff; vi = mt->cold_visc;
        co = mt->cold_cohesion;
    } else {
        nsm = mt->mass_mult; rd = mt->rest_dens; st = mt->stiff; vi = mt->visc; co = mt->cohesion;
        if (mt->cold_point != 0.0f) hot = true;
    }
    s.sph_mass[i] = nsm;
    s.m[i] = base_mass * nsm;
    s.rest_dens[i] = rd;
    s.stiff[i] = st;
    s.visc[i] = vi;
    s.cohesion[i] = co;
    if (mt->cold_point == 0.0f) {
        if (t <= mt->hot_point && hot) { solid = true; hot = false; }
    } else {
        if (t <= mt->cold_point && hot) { solid = true; hot = false; }
    }
    f &= ~(GE_FLAG_HOT_POINT | GE_FLAG_SOLIDIFIED);
    s.flags[i] = f | (hot ? GE_FLAG_HOT_POINT : 0u) | (solid ? GE_FLAG_SOLIDIFIED : 0u);
}

int fail(ge_ctx *c, int code, const char *msg) {
    c->err = msg;
    return code;
}

}  // namespace

void launch_thermal_fill(ge_ctx *c, const State &s, uint32_t at, uint32_t cnt) {
    if (cnt == 0) return;
    hipLaunchKernelGGL(k_thermal_fill, dim3(blocks_for(cnt, kBlock)), dim3(kBlock), 0, c->stream, s.temp, s.label,
                       s.ke, s.prev_ke, at, cnt);
}

// The average temperature of each node, gridTemp / particle count, as the
// walk's heat transfer divides it (physics.cpp:261-276): once per node here
// instead of once per accepting wave and node.
__global__ __launch_bounds__(kBlock) void k_node_tavg(const NodeRec *__restrict__ nodes,
                                                      const uint32_t *__restrict__ counters,
                                                      const float *__restrict__ node_temp,
                                                      float *__restrict__ tavg) {
    const uint32_t nn = node_total(counters);
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nn) return;
    const uint32_t cn = nodes[k].end - nodes[k].start;
    tavg[k] = cn > 0u ? node_temp[k] / (float)cn : 0.0f;
}

void launch_node_temps(ge_ctx *c) {
    const State &s = c->b.st[c->b.cur];
    const uint32_t cap = c->b.node_cap;
    const uint32_t nb = blocks_for(cap, kBlock);
    hipLaunchKernelGGL(k_node_links_init, dim3(nb), dim3(kBlock), 0, c->stream, c->b.node_parent, c->b.node_arrive,
                       cap);
    hipLaunchKernelGGL(k_node_links, dim3(nb), dim3(kBlock), 0, c->stream, c->b.nodes, c->b.counters,
                       c->b.node_parent);
    hipLaunchKernelGGL(k_node_temps_up, dim3(nb), dim3(kBlock), 0, c->stream, c->b.nodes, c->b.counters,
                       c->b.node_parent, c->b.node_arrive, s.temp, c->n, c->b.node_temp);
    hipLaunchKernelGGL(k_node_tavg, dim3(nb), dim3(kBlock), 0, c->stream, c->b.nodes, c->b.counters,
                       c->b.node_temp, c->b.node_tavg);
}

void launch_temperature(ge_ctx *c, const ge_step_params *p) {
    const State &s = c->b.st[c->b.cur];
    const ge_thermal_params &t = c->thermal_p;
    hipLaunchKernelGGL(k_temperature, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s, c->n,
                       c->b.materials, c->nmat, t.ambient_temp, t.ambient_heat_rate, p->dt, t.base_mass);
}

}  // namespace ge

using namespace ge;

namespace {

// materialsSPH.h:49-300, in SPHMaterials::Init's order
ge_material material(uint32_t id, float mass, float rest, float stiff, float visc, float coh, float hot, float hrest,
                     float hmass, float hstiff, float hvisc, float hcoh, float cold, float crest, float cmass,
                     float cstiff, float cvisc, float ccoh, float k, float res, float pp_frac, float ppm, float cst,
                     int plastic) {
    ge_material m;
    m.id = id;
    m.mass_mult = mass; m.rest_dens = rest; m.stiff = stiff; m.visc = visc; m.cohesion = coh;
    m.hot_point = hot; m.hot_rest_dens = hrest; m.hot_mass_mult = hmass; m.hot_stiff = hstiff;
    m.hot_visc = hvisc; m.hot_cohesion = hcoh;
    m.cold_point = cold; m.cold_rest_dens = crest; m.cold_mass_mult = cmass; m.cold_stiff = cstiff;
    m.cold_visc = cvisc; m.cold_cohesion = ccoh;
    m.heat_conductivity = k;
    m.constraint_resistance = res;
    m.constraint_plastic_point = res * pp_frac;  // constraintPlasticPoint = constraintResistance * f
    m.constraint_plastic_point_mult = ppm;
    m.constraint_stiffness = cst;
    m.is_plastic = plastic;
    return m;
}

}  // namespace

extern "C" {

void ge_default_thermal_params(ge_thermal_params *t) {
    if (!t) return;
    std::memset(t, 0, sizeof *t);
    t->heat_conductivity = 0.045f;  // parameters.h:120-122
    t->ambient_temp = 274.0f;
    t->ambient_heat_rate = 1.0f;
    t->base_mass = 8500000000.0f;   // parameters.cpp:44
}

int ge_thermal_enabled(const ge_ctx *c, int32_t *on) {
    if (!c || !on) return GE_E_INVALID;
    *on = c->thermal ? 1 : 0;
    return GE_OK;
}

int ge_default_materials(ge_material *out, uint32_t cap, uint32_t *count) {
    if (!count || (cap && !out)) return GE_E_INVALID;
    // water, rock, iron, sand, soil, mud, rubber; a material with cold point 0
    // has its base values as cold values
    const ge_material m[7] = {
        material(1, 0.6f, 0.095f, 1.0f, 0.075f, 0.05f, 373.2f, 0.065f, 0.25f, 1.0f, 0.055f, 0.0f, 273.2f, 0.046f,
                 0.5f, 0.6f, 2.2f, 2500.0f, 0.15f, 0.06f, 0.5f, 0.0f, 55.0f, 0),
        material(2, 3.3f, 0.008f, 1.4f, 3.0f, 1750.0f, 1370.0f, 0.005f, 2.8f, 1.0f, 0.6f, 300.0f, 0.0f, 0.008f,
                 3.3f, 1.4f, 3.0f, 1750.0f, 0.02f, 0.211f, 0.5f, 0.0f, 60.0f, 0),
        material(3, 4.0f, 0.008f, 1.4f, 3.0f, 1750.0f, 1740.0f, 0.005f, 2.8f, 1.0f, 0.6f, 300.0f, 0.0f, 0.008f,
                 4.0f, 1.4f, 3.0f, 1750.0f, 0.02f, 0.24f, 0.45f, 1.8f, 60.0f, 1),
        material(4, 2.1f, 0.008f, 1.255f, 0.74f, 1.0f, 1200.0f, 0.011f, 1.9f, 1.12f, 0.6f, 1.0f, 0.0f, 0.008f,
                 2.1f, 1.255f, 0.74f, 1.0f, 0.01f, 0.08f, 0.5f, 0.0f, 55.0f, 0),
        material(5, 1.9f, 0.008f, 1.0f, 2.23f, 3000.0f, 950.0f, 0.013f, 1.8f, 0.9f, 1.8f, 600.0f, 0.0f, 0.008f,
                 1.9f, 1.0f, 2.23f, 3000.0f, 0.02f, 0.09f, 0.5f, 2.0f, 30.0f, 1),
        material(6, 2.3f, 0.0095f, 1.0f, 0.6f, 100.0f, 1000.0f, 0.011f, 2.1f, 1.0f, 0.5f, 40.0f, 0.0f, 0.0095f,
                 2.3f, 1.0f, 0.6f, 100.0f, 0.1f, 0.03f, 0.5f, 2.0f, 20.0f, 1),
        material(7, 1.7f, 0.0095f, 1.0f, 0.6f, 100.0f, 453.0f, 0.011f, 1.6f, 1.0f, 0.5f, 40.0f, 0.0f, 0.0095f,
                 1.7f, 1.0f, 0.6f, 100.0f, 1.1f, 5.5f, 0.5f, 2.0f, 6.0f, 1),
    };
    *count = 7;
    for (uint32_t k = 0; k < 7 && k < cap; ++k) out[k] = m[k];
    return GE_OK;
}

int ge_set_thermal(ge_ctx *c, const ge_thermal_params *t, const ge_material *mats, uint32_t count) {
    if (!c || !t || (count && !mats)) return GE_E_INVALID;
    if (!c->thermal) return fail(c, GE_E_INVALID, "context created without enable_temperature");
    if (count > GE_MAX_MATERIALS) return fail(c, GE_E_INVALID, "more than GE_MAX_MATERIALS materials");
    (void)hipSetDevice(c->device);
    if (count) {
        if (hipMemcpy(c->b.materials, mats, sizeof(ge_material) * count, hipMemcpyHostToDevice) != hipSuccess)
            return fail(c, GE_E_HIP, "material copy failed");
    }
    c->thermal_p = *t;
    c->nmat = count;
    return GE_OK;
}

int ge_thermal_upload(ge_ctx *c, const ge_thermal_fields *f) {
    if (!c || !f) return GE_E_INVALID;
    if (!c->thermal) return fail(c, GE_E_INVALID, "context created without enable_temperature");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    const State &s = c->b.st[c->b.cur];
    const size_t b = sizeof(float) * c->n;
    if ((f->temp && hipMemcpyAsync(s.temp, f->temp, b, hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
        (f->label && hipMemcpyAsync(s.label, f->label, b, hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
        (f->ke && hipMemcpyAsync(s.ke, f->ke, b, hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
        (f->prev_ke && hipMemcpyAsync(s.prev_ke, f->prev_ke, b, hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(c, GE_E_HIP, "temperature field copy failed");
    return GE_OK;
}

int ge_thermal_download(ge_ctx *c, const ge_thermal_fields *f) {
    if (!c || !f) return GE_E_INVALID;
    if (!c->thermal) return fail(c, GE_E_INVALID, "context created without enable_temperature");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    const State &s = c->b.st[c->b.cur];
    const size_t b = sizeof(float) * c->n;
    if ((f->temp && hipMemcpyAsync(f->temp, s.temp, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        (f->label && hipMemcpyAsync(f->label, s.label, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        (f->ke && hipMemcpyAsync(f->ke, s.ke, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        (f->prev_ke && hipMemcpyAsync(f->prev_ke, s.prev_ke, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(c, GE_E_HIP, "temperature field copy failed");
    return GE_OK;
}

int ge_temperature_calculation(ge_ctx *c, const ge_step_params *p) {
    if (!c || !p) return GE_E_INVALID;
    if (!c->thermal) return fail(c, GE_E_INVALID, "context created without enable_temperature");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch_temperature(c, p);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, GE_E_HIP, "temperature pass failed");
    return GE_OK;
}

int ge_get_node_temps(ge_ctx *c, float *out, uint32_t cap) {
    if (!c || (cap && !out)) return GE_E_INVALID;
    if (!c->thermal) return fail(c, GE_E_INVALID, "context created without enable_temperature");
    if (!c->have_tree) return fail(c, GE_E_INVALID, "no tree built");
    (void)hipSetDevice(c->device);
    launch_node_temps(c);
    uint32_t total = 0;
    if (hipMemcpyAsync(&total, c->b.counters + kCntNodeTotal, sizeof total, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(c, GE_E_HIP, "node count copy failed");
    const uint32_t m = total < cap ? total : cap;
    if (m && (hipMemcpyAsync(out, c->b.node_temp, sizeof(float) * m, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
        return fail(c, GE_E_HIP, "node temperature copy failed");
    return GE_OK;
}

}  // extern "C"
