// ge_sph3d.hip -- SPH3D on the device: the neighbour grid of
// NeighborSearch3DV2AVX2 (GE/include/Particles/neighborSearch.h:978-1193) and
// the PCISPH passes of SPH3D::pcisphSolver (GE/include/Physics/SPH3D.h:
// 590-599, GE/src/Physics/SPH3D.cpp:98-366).
//
// Candidates of a query point q are every particle whose cell of pos,
// floor(p / cellSize) per axis, lies in the 3 x 3 x 3 block around q's cell:
// no distance filter, self included (queryNeighborsAVX2,
// neighborSearch.h:1109-1173).  The reference groups cells by a 32768-slot
// hash and compares exact coordinates inside a slot; here particles are
// radix-sorted by a slot of a 2^k-slot hash (k sized to the context, so a slot
// rarely holds more than one cell), first[] gives each slot's range, and the
// exact-coordinate test makes the candidate SET the reference's whatever the
// hash.  Cell coordinates and the pack of each candidate's fields are stored
// in slot order, so a slot is read contiguously.
//
// The reference scatters every pair force to both particles with OpenMP
// atomics.  Here each particle gathers both halves itself (no float atomics;
// deterministic), as the 2D passes do:
//   viscosity/cohesion (queries at pos, a symmetric relation): i's own visits
//     (+F(i, j)) and the visits of i by its candidates (-F(j, i)) from one
//     candidate set;
//   pressure (queries at predPos, not symmetric): own visits from grid A
//     around cell(predPos_i); visits of i by others from grid B (cells of the
//     predicted positions of visiting particles) around cell(pos_i).
// Kernel constants follow SPH3D.h:27-53 in float arithmetic (pow with float
// arguments is std::pow(float, float) in the reference's build).
#include <algorithm>
#include <cmath>
#include <vector>

#include "ge_internal3d.hpp"

namespace ge {
namespace {

constexpr float kPiF = 3.14159265358979323846f;  // raylib PI, a float

struct Sph3Const {
    float h, h2, cell, unit, unit2, visc, coh, delta, stiff_mult, coeff;
    float vol, spiky, lap;  // SPH3D.h:27-46
    float half_dt;          // fused integrateEnd3D (0 = none)
    uint32_t mask;          // hash slots - 1
    float h2hi;             // a bound with sqrt(r2) < h  =>  r2 < h2hi (cheap pre-test)
    float inv_h, inv_vol, coh_norm;  // 1/h, 1/vol, 30/(pi h^3)
};

// The pair passes divide with the hardware reciprocal (v_rcp_f32, within
// 1 ulp) where the reference divides: the pair terms then differ from the
// reference's by a few ulp, far inside the passes' tolerance.  Tests that
// decide which pairs interact keep the reference's exact arithmetic.
__device__ __forceinline__ float rcp3(float v) { return __builtin_amdgcn_rcpf(v); }

struct Grid3View {
    const uint32_t *first;
    const uint32_t *ord;
    const int4 *cell;
    const float4 *pk0, *pk1;
    const float2 *pk2;
};

__device__ __forceinline__ bool visitor3(uint32_t f) { return (f & GE_FLAG_SPH) && !(f & GE_FLAG_DRAWN); }

// posToCellCoord (neighborSearch.h:1005-1010), clamped so that a far-away
// point cannot overflow the int conversion
__device__ __forceinline__ int cell_of(float p, float cell) {
    float c = floorf(p / cell);
    c = fminf(fmaxf(c, -1073741824.0f), 1073741824.0f);
    return (int)c;
}

__device__ __forceinline__ uint32_t slot_of(int cx, int cy, int cz, uint32_t mask) {
    uint32_t h = (uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^ (uint32_t)cz * 83492791u;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h & mask;
}

// Calls f(kk, flags) for every candidate position kk of the query cell
// (cx, cy, cz) in the reference's dz, dy, dx order.
template <class F>
__device__ __forceinline__ void for_candidates(const Grid3View &G, int cx, int cy, int cz, uint32_t mask, F &&f) {
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
                const uint32_t s = slot_of(nx, ny, nz, mask);
                const uint32_t lo = G.first[s], hi = G.first[s + 1];
                for (uint32_t kk = lo; kk < hi; ++kk) {
                    const int4 c = G.cell[kk];
                    if (c.x == nx && c.y == ny && c.z == nz) f(kk, (uint32_t)c.w);
                }
            }
}

// SPH3D::smoothingKernelCohesion (SPH3D.h:48-53)
__device__ __forceinline__ float coh_kernel3(float r, const Sph3Const &K) {
#pragma clang fp contract(off)
    if (r >= K.h) return 0.0f;
    const float q = r * K.inv_h;
    return (1.0f - q) * (0.5f - q) * (0.5f - q) * K.coh_norm;
}

// packed sort words (slot << kIdxBits | particle); particles outside the grid
// (grid B: non-visitors) take slot mask + 1, past every table entry
__global__ __launch_bounds__(kBlock) void k3_slot_keys(const float *__restrict__ x, const float *__restrict__ y,
                                                       const float *__restrict__ z, const uint32_t *__restrict__ flags,
                                                       uint32_t n, int visitors_only, float cell, uint32_t mask,
                                                       uint64_t *__restrict__ keys) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t s = (uint64_t)mask + 1u;
    if (!visitors_only || visitor3(flags[i]))
        s = slot_of(cell_of(x[i], cell), cell_of(y[i], cell), cell_of(z[i], cell), mask);
    keys[i] = (s << kIdxBits) | i;
}

// first[s] = number of sorted words whose slot is below s, for s in [0, mask + 1]
__global__ __launch_bounds__(kBlock) void k3_slot_table(const uint64_t *__restrict__ sorted, uint32_t n,
                                                        uint32_t nslots, uint32_t *__restrict__ first) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s > nslots) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((sorted[mid] >> kIdxBits) < s) lo = mid + 1;
        else hi = mid;
    }
    first[s] = lo;
}

// grid A packs in slot order: cell and flags, (x, y, z, sphMass),
// (vx, vy, vz, visc), (cohesion, dens)
__global__ __launch_bounds__(kBlock) void k3_pack_a(State3 s, uint32_t n, const uint64_t *__restrict__ sorted, float cell,
                                                    uint32_t *__restrict__ ord, int4 *__restrict__ cl,
                                                    float4 *__restrict__ pk0, float4 *__restrict__ pk1,
                                                    float2 *__restrict__ pk2) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = (uint32_t)(sorted[k] & kIdxMask);
    ord[k] = i;
    const float x = s.x[i], y = s.y[i], z = s.z[i];
    cl[k] = make_int4(cell_of(x, cell), cell_of(y, cell), cell_of(z, cell), (int)s.flags[i]);
    pk0[k] = make_float4(x, y, z, s.sph_mass[i]);
    pk1[k] = make_float4(s.vx[i], s.vy[i], s.vz[i], s.visc[i]);
    pk2[k] = make_float2(s.cohesion[i], s.dens[i]);
}

// Neighbour lists of the density and pressure passes: a scan kernel walks a
// query's candidates once and keeps the positions that pass the pass's
// distance test (about 15 % of the 27 cells' candidates lie within h), so the
// pair kernels that follow loop over dense lists and a wave's lanes do the
// pair work together.  (The viscosity pass reads each candidate set once and
// gains nothing from a list.)
// Column-major: entry t of sorted position k at idx[t * n + k].  A query
// with more than `cap` kept candidates scans its candidates again instead;
// both paths visit the same positions in the same (reference) order.
struct NbrList {
    uint32_t *idx;
    uint32_t *cnt;
    uint32_t cap;
};

template <class Keep>
__device__ __forceinline__ void store_kept(const NbrList &L, uint32_t k, uint32_t n, const Grid3View &G, int cx,
                                           int cy, int cz, uint32_t mask, Keep &&keep) {
    uint32_t m = 0;
    for_candidates(G, cx, cy, cz, mask, [&](uint32_t kk, uint32_t fj) {
        if (!keep(kk, fj)) return;
        if (m < L.cap) L.idx[(size_t)m * n + k] = kk;
        ++m;
    });
    L.cnt[k] = m;
}

template <class Keep, class Body>
__device__ __forceinline__ void for_kept(const NbrList &L, uint32_t k, uint32_t n, const Grid3View &G, int cx, int cy,
                                         int cz, uint32_t mask, Keep &&keep, Body &&body) {
    const uint32_t m = L.cnt[k];
    if (m <= L.cap) {
        for (uint32_t t = 0; t < m; ++t) body(L.idx[(size_t)t * n + k]);
    } else {
        for_candidates(G, cx, cy, cz, mask, [&](uint32_t kk, uint32_t fj) {
            if (keep(kk, fj)) body(kk);
        });
    }
}

// the density pass's test (SPH3D.cpp:208-222): a visiting particle whose
// predicted position is within h of the query's (self included)
struct PredKeep {
    const float4 *pp;
    float4 me;
    float h;
    __device__ __forceinline__ bool operator()(uint32_t kk, uint32_t fj) const {
#pragma clang fp contract(off)
        if (!visitor3(fj)) return false;
        const float4 pj = pp[kk];
        const float dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;
        return dx * dx + dy * dy + dz * dz < h * h;
    }
};

// the reverse pressure visits (SPH3D.cpp:244-256 from the visitor's side):
// another visitor whose predicted position is in [1e-5, h) of i's
struct RevKeep {
    const Grid3View &B;
    uint32_t i;
    float4 me;
    float h;
    __device__ __forceinline__ bool operator()(uint32_t kb, uint32_t) const {
#pragma clang fp contract(off)
        if (B.ord[kb] == i) return false;
        const float4 pj = B.pk0[kb];
        const float dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;
        const float rr = sqrtf(dx * dx + dy * dy + dz * dz);
        return !(rr < 1e-5f || rr >= h);
    }
};

// The viscosity pass gives each particle kSph3Lanes lanes: lane q takes the
// cells q, q + kSph3Lanes, ... of the 27 (numbered in the reference's dz, dy,
// dx order), so the cells' hashes and slot ranges are read in parallel, and
// the group adds its partial sums in a fixed butterfly order (deterministic).
// (The density and pressure passes keep one thread per particle over the
// neighbour lists below: eight lanes scanning the 27 cells twice per
// pressure pass measured slower than building the lists once.)
constexpr uint32_t kSph3Lanes = 8;

__device__ __forceinline__ float group_sum3(float v) {
#pragma unroll
    for (uint32_t o = 1; o < kSph3Lanes; o <<= 1) v += __shfl_xor(v, (int)o, 64);
    return v;
}

// f(kk, flags) for the candidates of this lane's cells around (cx, cy, cz)
template <class F>
__device__ __forceinline__ void for_lane_candidates(const Grid3View &G, int cx, int cy, int cz, uint32_t mask,
                                                    uint32_t q0, F &&f) {
    for (uint32_t c = q0; c < 27u; c += kSph3Lanes) {
        const int nx = cx + (int)(c % 3u) - 1, ny = cy + (int)((c / 3u) % 3u) - 1, nz = cz + (int)(c / 9u) - 1;
        const uint32_t s = slot_of(nx, ny, nz, mask);
        const uint32_t lo = G.first[s], hi = G.first[s + 1];
        for (uint32_t kk = lo; kk < hi; ++kk) {
            const int4 cc = G.cell[kk];
            if (cc.x == nx && cc.y == ny && cc.z == nz) f(kk, (uint32_t)cc.w);
        }
    }
}

// SPH3D::computeViscCohesionForces (SPH3D.cpp:98-170) in gather form, then
// the predicted position (SPH3D.cpp:190-198).  pp[k] = (ppx, ppy, ppz, restDens).
__global__ __launch_bounds__(kBlock) void k3_visc(State3 s, uint32_t n, Grid3View A, Sph3Const K,
                                                  float4 *__restrict__ sphf, float4 *__restrict__ pp) {
#pragma clang fp contract(off)
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t k = t / kSph3Lanes, q0 = t % kSph3Lanes;
    if (k >= n) return;
    const uint32_t i = A.ord[k];
    const int4 ci = A.cell[k];
    const uint32_t fi = (uint32_t)ci.w;
    if (!visitor3(fi)) {
        if (q0 == 0) {
            sphf[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            pp[k] = make_float4(s.ppx[i], s.ppy[i], s.ppz[i], s.rest_dens[i]);
        }
        return;
    }
    const float4 o0 = A.pk0[k], o1 = A.pk1[k];
    const float2 o2 = A.pk2[k];
    const bool fwd = !(fi & GE_FLAG_PINNED);
    const float mI = o0.w * K.unit;
    const float di = o2.y < 0.001f ? 0.001f : o2.y;
    const float aI = K.visc * o1.w * mI * rcp3(di);  // j's visit of i: viscosity factor without the kernel
    const float cI = K.coh * o2.x;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for_lane_candidates(A, ci.x, ci.y, ci.z, K.mask, q0, [&](uint32_t kk, uint32_t fj) {
        if (kk == k || !visitor3(fj)) return;
        const float4 p0 = A.pk0[kk];
        const float dx = p0.x - o0.x, dy = p0.y - o0.y, dz = p0.z - o0.z;
        const float rsq = dx * dx + dy * dy + dz * dz;
        if (rsq >= K.h2) return;
        const float r = __builtin_amdgcn_sqrtf(rsq < 1e-6f ? 1e-6f : rsq);
        const float inv_r = rcp3(r);
        const float lapw = r >= K.h ? 0.0f : K.lap * (K.h - r);  // smoothingKernelLaplacian
        const float wc = coh_kernel3(r, K);
        const float4 p1 = A.pk1[kk];
        const float2 p2 = A.pk2[kk];
        if (fwd) {  // i's visit of j (SPH3D.cpp:108-158)
            const float mJ = p0.w * K.unit;
            const float dj = p2.y < 0.001f ? 0.001f : p2.y;
            const float a = K.visc * p1.w * mJ * rcp3(dj) * lapw;
            const float vfx = a * (p1.x - o1.x), vfy = a * (p1.y - o1.y), vfz = a * (p1.z - o1.z);
            const float cw = cI * mJ * wc * inv_r;
            const float cfx = cw * dx, cfy = cw * dy, cfz = cw * dz;
            sx = sx + (vfx + cfx);
            sy = sy + (vfy + cfy);
            sz = sz + (vfz + cfz);
        }
        if (!(fj & GE_FLAG_PINNED)) {  // j's visit of i, scattered back to i (SPH3D.cpp:160-168)
            const float a = aI * lapw;
            const float vfx = a * (o1.x - p1.x), vfy = a * (o1.y - p1.y), vfz = a * (o1.z - p1.z);
            const float cw = K.coh * p2.x * mI * wc * inv_r;
            const float cfx = cw * -dx, cfy = cw * -dy, cfz = cw * -dz;  // pos_i - pos_j
            sx = sx - (vfx + cfx);
            sy = sy - (vfy + cfy);
            sz = sz - (vfz + cfz);
        }
    });
    sx = group_sum3(sx);
    sy = group_sum3(sy);
    sz = group_sum3(sz);
    if (q0 != 0) return;
    sphf[i] = make_float4(sx, sy, sz, 0.0f);
    const float px = o0.x + (sx / o0.w) * K.coeff;
    const float py = o0.y + (sy / o0.w) * K.coeff;
    const float pz = o0.z + (sz / o0.w) * K.coeff;
    s.ppx[i] = px;
    s.ppy[i] = py;
    s.ppz[i] = pz;
    pp[k] = make_float4(px, py, pz, s.rest_dens[i]);
}

__global__ __launch_bounds__(kBlock) void k3_pred_scan(uint32_t n, Grid3View A, const float4 *__restrict__ pp,
                                                       Sph3Const K, NbrList L) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    if (!visitor3((uint32_t)A.cell[k].w)) {
        L.cnt[k] = 0;
        return;
    }
    const float4 me = pp[k];
    store_kept(L, k, n, A, cell_of(me.x, K.cell), cell_of(me.y, K.cell), cell_of(me.z, K.cell), K.mask,
               PredKeep{pp, me, K.h});
}

// predicted density and pressure (SPH3D.cpp:200-236); pd[k] = (press, predDens)
__global__ __launch_bounds__(kBlock) void k3_density(State3 s, uint32_t n, Grid3View A, const float4 *__restrict__ pp,
                                                     Sph3Const K, NbrList L, float2 *__restrict__ pd) {
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = A.ord[k];
    const uint32_t fi = (uint32_t)A.cell[k].w;
    if (!visitor3(fi)) {
        s.press[i] = 0.0f;  // press is cleared for every particle (SPH3D.cpp:182-185)
        pd[k] = make_float2(0.0f, s.pred_dens[i]);
        return;
    }
    const float4 me = pp[k];
    float dens = 0.0f;
    for_kept(L, k, n, A, cell_of(me.x, K.cell), cell_of(me.y, K.cell), cell_of(me.z, K.cell), K.mask,
             PredKeep{pp, me, K.h}, [&](uint32_t kk) {
                 const float4 pj = pp[kk];
                 const float dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;
                 const float rr = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
                 const float mJ = A.pk0[kk].w * K.unit2;
                 const float rho0 = 0.5f * (me.w + pj.w);
                 const float w = rr >= K.h ? 0.0f : (K.h - rr) * (K.h - rr) * K.inv_vol;  // smoothingKernel
                 dens = dens + mJ * w * rcp3(rho0);
             });
    float pt = K.delta * (dens - me.w);
    if (pt < 0.0f) pt = 0.0f;
    const float pr = 0.0f + pt * s.stiff[i] * K.stiff_mult;
    s.pred_dens[i] = dens;
    s.press_tmp[i] = pt;
    s.press[i] = pr;
    pd[k] = make_float2(pr, dens);
}

// grid B packs in slot order: cell of predPos, (ppx, ppy, ppz, press),
// (predDens, sphMass).  Non-visiting particles sort past first[nslots] and
// are never read.
__global__ __launch_bounds__(kBlock) void k3_pack_b(State3 s, uint32_t n, const uint64_t *__restrict__ sorted,
                                                    float cell, uint32_t *__restrict__ ord, int4 *__restrict__ cl,
                                                    float4 *__restrict__ pk0, float4 *__restrict__ pk1) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = (uint32_t)(sorted[k] & kIdxMask);
    ord[k] = i;
    const float x = s.ppx[i], y = s.ppy[i], z = s.ppz[i];
    cl[k] = make_int4(cell_of(x, cell), cell_of(y, cell), cell_of(z, cell), (int)s.flags[i]);
    pk0[k] = make_float4(x, y, z, s.press[i]);
    pk1[k] = make_float4(s.pred_dens[i], s.sph_mass[i], 0.0f, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k3_rev_scan(uint32_t n, Grid3View A, const float4 *__restrict__ pp,
                                                      Grid3View B, Sph3Const K, NbrList L) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const int4 ci = A.cell[k];
    if (!visitor3((uint32_t)ci.w)) {
        L.cnt[k] = 0;
        return;
    }
    store_kept(L, k, n, B, ci.x, ci.y, ci.z, K.mask, RevKeep{B, A.ord[k], pp[k], K.h});
}

// pressure force of one visit (SPH3D.cpp:244-281); d = predPos_visitor - predPos_other
__device__ __forceinline__ void pressure_pair3(float dx, float dy, float dz, float rr, float mi, float mj, float pi_,
                                               float pj_, float di, float dj, const Sph3Const &K, float &fx, float &fy,
                                               float &fz) {
#pragma clang fp contract(off)
    const float gw = rr >= K.h ? 0.0f : K.spiky * ((K.h - rr) * (K.h - rr));  // spikyKernelDerivative
    const float inv_rr = rcp3(rr);
    const float nx = dx * inv_rr, ny = dy * inv_rr, nz = dz * inv_rr;
    const float avgP = 0.5f * (pi_ + pj_);
    const float avgD = 0.5f * (di + dj);
    float mag = -(mi * K.unit2 + mj * K.unit2) * avgP * rcp3(avgD < 0.01f ? 0.01f : avgD);
    const float smax = mi < mj ? mj : mi;
    const float smin = mj < mi ? mj : mi;
    const float sc8 = 8.0f * smin * rcp3(smax);  // 8 / (smax / smin)
    mag = mag * (sc8 < 1.0f ? sc8 : 1.0f);
    fx = mag * gw * nx;
    fy = mag * gw * ny;
    fz = mag * gw * nz;
}

// pressure forces (SPH3D.cpp:238-291), pressF and acc (SPH3D.cpp:296-307),
// then the integrateEnd3D kick (physics3D.cpp:1158-1168) when asked
__global__ __launch_bounds__(kBlock) void k3_pressure(State3 s, uint32_t n, Grid3View A, const float4 *__restrict__ pp,
                                                      const float2 *__restrict__ pd, Grid3View B, Sph3Const K,
                                                      NbrList Lp, NbrList Lr, const float4 *__restrict__ sphf) {
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = A.ord[k];
    const int4 ci = A.cell[k];
    const uint32_t fi = (uint32_t)ci.w;
    const float4 f0 = sphf[i];
    float fx = f0.x, fy = f0.y, fz = f0.z;
    const float smi = A.pk0[k].w;
    if (visitor3(fi)) {
        const float4 me = pp[k];
        const float2 mp = pd[k];
        // i's visits: the density pass's list (candidates by pos cell around
        // cell(predPos_i) within h), without i itself and closer than 1e-5
        for_kept(Lp, k, n, A, cell_of(me.x, K.cell), cell_of(me.y, K.cell), cell_of(me.z, K.cell), K.mask,
                 PredKeep{pp, me, K.h}, [&](uint32_t kk) {
                     if (kk == k) return;
                     const float4 pj = pp[kk];
                     const float dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;
                     const float rr = sqrtf(dx * dx + dy * dy + dz * dz);
                     if (rr < 1e-5f || rr >= K.h) return;
                     const float2 q = pd[kk];
                     float px, py, pz;
                     pressure_pair3(dx, dy, dz, rr, smi, A.pk0[kk].w, mp.x, q.x, mp.y, q.y, K, px, py, pz);
                     fx = fx + px;
                     fy = fy + py;
                     fz = fz + pz;
                 });
        // visits of i by others: predPos cells around cell(pos_i)
        for_kept(Lr, k, n, B, ci.x, ci.y, ci.z, K.mask, RevKeep{B, i, me, K.h}, [&](uint32_t kb) {
            const float4 pj = B.pk0[kb];
            const float dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;
            const float rr = sqrtf(dx * dx + dy * dy + dz * dz);
            const float4 q = B.pk1[kb];
            float px, py, pz;
            pressure_pair3(dx, dy, dz, rr, q.y, smi, pj.w, mp.x, q.x, mp.y, K, px, py, pz);
            fx = fx - px;
            fy = fy - py;
            fz = fz - pz;
        });
    }
    const float pfx = fx / smi, pfy = fy / smi, pfz = fz / smi;
    s.pfx[i] = pfx;
    s.pfy[i] = pfy;
    s.pfz[i] = pfz;
    if (!(fi & GE_FLAG_PINNED)) {
        const float ax = s.ax[i] + pfx, ay = s.ay[i] + pfy, az = s.az[i] + pfz;
        s.ax[i] = ax;
        s.ay[i] = ay;
        s.az[i] = az;
        if (K.half_dt != 0.0f) {
            s.vx[i] = s.vx[i] + ax * K.half_dt;
            s.vy[i] = s.vy[i] + ay * K.half_dt;
            s.vz[i] = s.vz[i] + az * K.half_dt;
        }
    }
}

// SPH3D::groundModeBoundary (SPH3D.cpp:310-366): vertical gravity and six
// damped walls at radiusMultiplier inside +-domain/2
__global__ __launch_bounds__(kBlock) void k3_ground(State3 s, uint32_t n, float radius, float w, float h, float d,
                                                    float vertical_gravity, float friction) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || (s.flags[i] & GE_FLAG_PINNED)) return;
    const float damp = -0.1f;  // SPH3D::boundDamping (SPH3D.h:14)
    float x = s.x[i], y = s.y[i], z = s.z[i], vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
    s.ay[i] = s.ay[i] - vertical_gravity;
    bool hx = false, hy = false, hz = false;
    if (x - radius < -w / 2.0f) { vx *= damp; x = -w / 2.0f + radius; hx = true; }
    if (x + radius > w / 2.0f) { vx *= damp; x = w / 2.0f - radius; hx = true; }
    if (y - radius < -h / 2.0f) { vy *= damp; y = -h / 2.0f + radius; hy = true; }
    if (y + radius > h / 2.0f) { vy *= damp; y = h / 2.0f - radius; hy = true; }
    if (z - radius < -d / 2.0f) { vz *= damp; z = -d / 2.0f + radius; hz = true; }
    if (z + radius > d / 2.0f) { vz *= damp; z = d / 2.0f - radius; hz = true; }
    if (hx) { vy *= 1.0f - friction; vz *= 1.0f - friction; }
    if (hy) { vx *= 1.0f - friction; vz *= 1.0f - friction; }
    if (hz) { vx *= 1.0f - friction; vy *= 1.0f - friction; }
    s.x[i] = x; s.y[i] = y; s.z[i] = z;
    s.vx[i] = vx; s.vy[i] = vy; s.vz[i] = vz;
}

// ---- candidate lists for the parity hook: counts, then lists; particles
// whose flags match skip_mask count 0
__global__ __launch_bounds__(kBlock) void k3_cand_count(const float *__restrict__ qx, const float *__restrict__ qy,
                                                        const float *__restrict__ qz, uint32_t n, Grid3View A,
                                                        float cell, uint32_t mask, const uint32_t *__restrict__ flags,
                                                        uint32_t skip_mask, int keep_skipped,
                                                        uint32_t *__restrict__ cnt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (flags && (flags[i] & skip_mask)) {
        if (!keep_skipped) cnt[i] = 0;
        return;
    }
    uint32_t c = 0;
    for_candidates(A, cell_of(qx[i], cell), cell_of(qy[i], cell), cell_of(qz[i], cell), mask,
                   [&](uint32_t, uint32_t) { ++c; });
    cnt[i] = c;
}

__global__ __launch_bounds__(kBlock) void k3_cand_fill(const float *__restrict__ qx, const float *__restrict__ qy,
                                                       const float *__restrict__ qz, uint32_t n, Grid3View A,
                                                       float cell, uint32_t mask, const int64_t *__restrict__ off,
                                                       uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int64_t w = off[i];
    for_candidates(A, cell_of(qx[i], cell), cell_of(qy[i], cell), cell_of(qz[i], cell), mask,
                   [&](uint32_t kk, uint32_t) { out[w++] = A.ord[kk]; });
}

Sph3Const make_const3(const ge3_ctx *c, const ge3_step_params *p, float half_dt) {
    Sph3Const K;
    K.h = p->sph_h;
    K.h2 = p->sph_h * p->sph_h;
    K.cell = p->sph_cell;
    K.unit = p->sph_unit_mass;
    K.unit2 = p->sph_unit_mass * 2.0f;  // myVar.mass * 2.0f (SPH3D.cpp:225, 262)
    K.visc = p->viscosity;
    K.coh = p->cohesion_coef;
    K.delta = p->delta;
    K.stiff_mult = p->stiff_mult;
    K.coeff = 0.5f * p->dt * p->dt;  // predictionCoeff (SPH3D.cpp:176)
    // SPH3D.h:27-46 in float
    K.vol = (2.0f * kPiF * std::pow(K.h, 5.0f)) / 15.0f;
    K.spiky = -45.0f / (kPiF * std::pow(K.h, 6.0f));
    K.lap = 45.0f / (kPiF * std::pow(K.h, 6.0f));
    K.half_dt = half_dt;
    K.mask = (1u << c->b.slot_bits) - 1u;
    // sqrt is correctly rounded and h a float, so sqrt(r2) < h implies r2 < h^2
    // exactly; the float bound sits a little above h^2
    K.h2hi = std::nextafter((float)((double)K.h * K.h * (1.0 + 1e-6)), INFINITY);
    K.inv_h = 1.0f / K.h;
    K.inv_vol = 1.0f / K.vol;
    K.coh_norm = 30.0f / (kPiF * K.h * K.h * K.h);
    return K;
}

Grid3View view3(const Grid3 &g) { return Grid3View{g.first, g.ord, g.cell, g.pk0, g.pk1, g.pk2}; }

// sort by slot and fill first[]
void build_slots(ge3_ctx *c, Grid3 &g) {
    const uint32_t n = c->n;
    const uint32_t nslots = 1u << c->b.slot_bits;
    radix_sort_words(c->stream, c->b.hist, g.keys, g.keys_alt, n, kIdxBits, kIdxBits + (int)c->b.slot_bits + 1,
                     c->b.counters + kCntError);
    hipLaunchKernelGGL(k3_slot_table, dim3(blocks_for((uint64_t)nslots + 1, kBlock)), dim3(kBlock), 0, c->stream,
                       g.keys, n, nslots, g.first);
}

}  // namespace

// NeighborSearch3DV2AVX2::newGridAVX2 (neighborSearch.h:1017-1107): grid A over pos
void sph3_new_grid(ge3_ctx *c, const ge3_step_params *p) {
    Grid3 &g = c->b.grid_pos;
    const State3 &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const uint32_t mask = (1u << c->b.slot_bits) - 1u;
    hipLaunchKernelGGL(k3_slot_keys, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, s.z, s.flags,
                       n, 0, p->sph_cell, mask, g.keys);
    build_slots(c, g);
    hipLaunchKernelGGL(k3_pack_a, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, s, n, g.keys, p->sph_cell,
                       g.ord, g.cell, g.pk0, g.pk1, g.pk2);
    c->have_grid = true;
    c->grid_cell = p->sph_cell;
}

// SPH3D::PCISPH (SPH3D.cpp:172-308), one iteration, on grid A
void sph3_pcisph(ge3_ctx *c, const ge3_step_params *p, float end_kick_dt) {
    Grid3 &A = c->b.grid_pos;
    Grid3 &B = c->b.grid_pred;
    const State3 &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const Sph3Const K = make_const3(c, p, end_kick_dt * 0.5f);
    const uint32_t nb = blocks_for(n, kBlock);
    const NbrList Lp{c->b.nbr_idx[0], c->b.nbr_cnt[0], c->b.nbr_cap};
    const NbrList Lr{c->b.nbr_idx[1], c->b.nbr_cnt[1], c->b.nbr_cap};
    // the viscosity pass: kSph3Lanes per particle
    hipLaunchKernelGGL(k3_visc, dim3(blocks_for((uint64_t)n * kSph3Lanes, kBlock)), dim3(kBlock), 0, c->stream, s, n,
                       view3(A), K, c->b.sphf, c->b.sph_pp);
    // grid B over the predicted positions of visiting particles
    hipLaunchKernelGGL(k3_slot_keys, dim3(nb), dim3(kBlock), 0, c->stream, s.ppx, s.ppy, s.ppz, s.flags, n, 1,
                       p->sph_cell, K.mask, B.keys);
    build_slots(c, B);
    hipLaunchKernelGGL(k3_pred_scan, dim3(nb), dim3(kBlock), 0, c->stream, n, view3(A), c->b.sph_pp, K, Lp);
    hipLaunchKernelGGL(k3_density, dim3(nb), dim3(kBlock), 0, c->stream, s, n, view3(A), c->b.sph_pp, K, Lp,
                       c->b.sph_pd);
    hipLaunchKernelGGL(k3_pack_b, dim3(nb), dim3(kBlock), 0, c->stream, s, n, B.keys, p->sph_cell, B.ord, B.cell,
                       B.pk0, B.pk1);
    hipLaunchKernelGGL(k3_rev_scan, dim3(nb), dim3(kBlock), 0, c->stream, n, view3(A), c->b.sph_pp, view3(B), K, Lr);
    hipLaunchKernelGGL(k3_pressure, dim3(nb), dim3(kBlock), 0, c->stream, s, n, view3(A), c->b.sph_pp, c->b.sph_pd,
                       view3(B), K, Lp, Lr, c->b.sphf);
}

void sph3_ground(ge3_ctx *c, const ge3_step_params *p) {
    hipLaunchKernelGGL(k3_ground, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.st[c->b.cur], c->n,
                       p->sph_h, p->domain_w, p->domain_h, p->domain_d, p->vertical_gravity, p->boundary_friction);
}

// queryNeighborsAVX2 candidate sets of grid A at pos (which 0) or predPos (1)
int sph3_neighbors(ge3_ctx *c, int which, int64_t *off, uint32_t *idx, int64_t cap, int64_t *total) {
    const Grid3 &g = c->b.grid_pos;
    const State3 &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const float *qx = which ? s.ppx : s.x, *qy = which ? s.ppy : s.y, *qz = which ? s.ppz : s.z;
    const uint32_t mask = (1u << c->b.slot_bits) - 1u;
    uint32_t *dcnt = reinterpret_cast<uint32_t *>(c->b.sphf);  // pair-force scratch: 4 floats per particle
    hipLaunchKernelGGL(k3_cand_count, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, qx, qy, qz, n, view3(g),
                       c->grid_cell, mask, nullptr, 0u, 0, dcnt);
    std::vector<uint32_t> hc(n);
    if (hipMemcpyAsync(hc.data(), dcnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "neighbour count copy failed";
        return GE_E_HIP;
    }
    off[0] = 0;
    for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + hc[i];
    *total = off[n];
    if (!idx) return GE_OK;
    if (cap < off[n]) {
        c->err = "neighbour buffer too small";
        return GE_E_CAPACITY;
    }
    int64_t *doff = nullptr;
    uint32_t *dout = nullptr;
    if (hipMalloc(&doff, sizeof(int64_t) * (n + 1)) != hipSuccess ||
        hipMalloc(&dout, sizeof(uint32_t) * (off[n] > 0 ? off[n] : 1)) != hipSuccess) {
        (void)hipFree(doff);
        c->err = "neighbour list allocation failed";
        return GE_E_NOMEM;
    }
    int rc = GE_OK;
    if (hipMemcpyAsync(doff, off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = GE_E_HIP;
    if (rc == GE_OK) {
        hipLaunchKernelGGL(k3_cand_fill, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, qx, qy, qz, n,
                           view3(g), c->grid_cell, mask, doff, dout);
        if (hipMemcpyAsync(idx, dout, sizeof(uint32_t) * off[n], hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = GE_E_HIP;
    }
    (void)hipFree(doff);
    (void)hipFree(dout);
    if (rc != GE_OK) {
        c->err = "neighbour list copy failed";
        return rc;
    }
    for (uint32_t i = 0; i < n; ++i) std::sort(idx + off[i], idx + off[i + 1]);
    return GE_OK;
}

// NeighborSearch3DV2AVX2::neighborAmount (neighborSearch.h:1175-1193) on grid A
int sph3_neighbor_counts(ge3_ctx *c, uint32_t *counts) {
    const Grid3 &g = c->b.grid_pos;
    const State3 &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const uint32_t mask = (1u << c->b.slot_bits) - 1u;
    uint32_t *dcnt = reinterpret_cast<uint32_t *>(c->b.sphf);
    hipLaunchKernelGGL(k3_cand_count, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, s.z, n,
                       view3(g), c->grid_cell, mask, s.flags, GE_FLAG_PINNED | GE_FLAG_DRAWN | GE_FLAG_DARK, 0, dcnt);
    if (hipMemcpyAsync(counts, dcnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "neighbour count copy failed";
        return GE_E_HIP;
    }
    return GE_OK;
}

// NeighborSearch3DV2AVX2::neighborAmount (neighborSearch.h:1175-1193) into
// out[] (device), leaving the counts of pinned, drawn and dark-matter
// particles as they were.  Rebuilds grid A with this cell size.
void sph3_count_into(ge3_ctx *c, float cell, int32_t *out) {
    ge3_step_params q{};
    q.sph_cell = cell;
    sph3_new_grid(c, &q);
    const Grid3 &g = c->b.grid_pos;
    const State3 &s = c->b.st[c->b.cur];
    const uint32_t mask = (1u << c->b.slot_bits) - 1u;
    hipLaunchKernelGGL(k3_cand_count, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, s.z, c->n,
                       view3(g), cell, mask, s.flags, GE_FLAG_PINNED | GE_FLAG_DRAWN | GE_FLAG_DARK, 1,
                       reinterpret_cast<uint32_t *>(out));
    c->have_grid = false;
}

}  // namespace ge
