// ge_step3d.hip -- per-frame streaming kernels of the 3D step:
// integrateStart3D (with the SPH speed clamp) fused with the boundingBox3D
// reduction, Morton3D keys, the
// key-value sort and Morton-order gather, the wave-coherent octree walk of
// calculateForceFromGrid3D (with integrateEnd3D's kick fused in), the direct
// sum of naiveGravity3D, integrateEnd3D, and an FP64 direct sum for error
// checks.
#include <cfloat>

#include "ge_internal3d.hpp"

namespace ge {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ float wmin(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = glm_min(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ float wmax(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = glm_max(v, __shfl_xor(v, o, kWave));
    return v;
}

// workgroup min of mn[0..2] and max of mx[0..2]; result valid in thread 0
__device__ __forceinline__ void block_minmax3(float (&mn)[3], float (&mx)[3]) {
    __shared__ float red[6][kBlock / kWave];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        mn[k] = wmin(mn[k]);
        mx[k] = wmax(mx[k]);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            red[k][w] = mn[k];
            red[3 + k][w] = mx[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < kBlock / kWave; ++j) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                mn[k] = glm_min(mn[k], red[k][j]);
                mx[k] = glm_max(mx[k], red[3 + k][j]);
            }
        }
    }
}

// Physics3D::integrateStart3D (physics3D.cpp:740-804), then the per-workgroup
// part of boundingBox3D (globalLogic.cpp:1465-1477) over the drifted positions.
__global__ __launch_bounds__(kBlock) void k3_integrate_bbox(State3 s, uint32_t n, float dt, int do_integrate,
                                                            Domain3 dom, int sph_clamp, float max_vel,
                                                            float *__restrict__ partials) {
#pragma clang fp contract(off)
    const float half_dt = dt * 0.5f;
    const float max_vsq = max_vel * max_vel;
    const bool per = dom.w > 0.0f;
    const float hw = dom.w * 0.5f, hh = dom.h * 0.5f, hd = dom.d * 0.5f;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    const uint32_t lo = blockIdx.x * kBBoxChunk3;
    const uint32_t hi = min(n, lo + kBBoxChunk3);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        float x = s.x[i], y = s.y[i], z = s.z[i];
        if (do_integrate && !(s.flags[i] & GE_FLAG_PINNED)) {
            float pvx = s.vx[i], pvy = s.vy[i], pvz = s.vz[i];
            float vx = pvx + s.ax[i] * half_dt;
            float vy = pvy + s.ay[i] * half_dt;
            float vz = pvz + s.az[i] * half_dt;
            if (sph_clamp) {  // SPH speed clamp, physics3D.cpp:761-781
                const float vsq = vx * vx + vy * vy + vz * vz;
                if (vsq > max_vsq) {
                    const float psq = pvx * pvx + pvy * pvy + pvz * pvz;
                    if (psq > 0.00001f) {
                        const float inv = max_vel / sqrtf(psq);
                        pvx = pvx * inv; pvy = pvy * inv; pvz = pvz * inv;
                    }
                    const float inv = max_vel / sqrtf(vsq);
                    vx = vx * inv; vy = vy * inv; vz = vz * inv;
                }
            }
            s.pvx[i] = pvx; s.pvy[i] = pvy; s.pvz[i] = pvz;
            s.vx[i] = vx; s.vy[i] = vy; s.vz[i] = vz;
            x = x + vx * dt;
            y = y + vy * dt;
            z = z + vz * dt;
            if (per) {  // physics3D.cpp:782-800: the cube [-d/2, d/2)
                if (x < -hw) x += dom.w;
                else if (x >= hw) x -= dom.w;
                if (y < -hh) y += dom.h;
                else if (y >= hh) y -= dom.h;
                if (z < -hd) z += dom.d;
                else if (z >= hd) z -= dom.d;
            }
            s.x[i] = x; s.y[i] = y; s.z[i] = z;
        }
        mn[0] = glm_min(mn[0], x); mn[1] = glm_min(mn[1], y); mn[2] = glm_min(mn[2], z);
        mx[0] = glm_max(mx[0], x); mx[1] = glm_max(mx[1], y); mx[2] = glm_max(mx[2], z);
    }
    block_minmax3(mn, mx);
    if (threadIdx.x == 0) {
        reinterpret_cast<float4 *>(partials)[2 * blockIdx.x] = make_float4(mn[0], mn[1], mn[2], 0.0f);
        reinterpret_cast<float4 *>(partials)[2 * blockIdx.x + 1] = make_float4(mx[0], mx[1], mx[2], 0.0f);
    }
}

// one workgroup: the partials -> {x, y, z, size} (globalLogic.cpp:1472-1476)
__global__ __launch_bounds__(kBlock) void k3_bbox_final(const float *__restrict__ partials, uint32_t nparts,
                                                        float *__restrict__ bb) {
#pragma clang fp contract(off)
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t k = threadIdx.x; k < nparts; k += kBlock) {
        const float4 a = reinterpret_cast<const float4 *>(partials)[2 * k];
        const float4 b = reinterpret_cast<const float4 *>(partials)[2 * k + 1];
        mn[0] = glm_min(mn[0], a.x); mn[1] = glm_min(mn[1], a.y); mn[2] = glm_min(mn[2], a.z);
        mx[0] = glm_max(mx[0], b.x); mx[1] = glm_max(mx[1], b.y); mx[2] = glm_max(mx[2], b.z);
    }
    block_minmax3(mn, mx);
    if (threadIdx.x == 0) {
        const float size = glm_max(glm_max(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
        const float half = size * 0.5f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bb[k] = (mn[k] + mx[k]) * 0.5f - half;
            bb[4 + k] = mn[k];
            bb[8 + k] = mx[k];
        }
        bb[3] = size;
        bb[7] = 0.0f;
        bb[11] = 0.0f;
    }
}

// Morton::scaleToGrid3D (morton.cpp:67-75)
__device__ __forceinline__ uint64_t scale_to_grid3(float pos, float lo, float hi) {
#pragma clang fp contract(off)
    if (hi <= lo) return 0;
    const float c = (pos < lo) ? lo : ((hi < pos) ? hi : pos);
    const float norm = (c - lo) / (hi - lo);
    const uint64_t s = (uint64_t)(norm * 2097151.0f);
    return s < 2097151u ? s : 2097151u;
}

// bit i of a 21-bit value to bit 3i: the result of spreadBits3D's loop
// (morton.cpp:77-86) in five mask steps
__device__ __forceinline__ uint64_t spread3(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Morton::computeMortonKeys3D (morton.cpp:94-113): the box without a
// minimum extent; key per slot and the (key, slot) pair to sort
__global__ __launch_bounds__(kBlock) void k3_morton_keys(const float *__restrict__ x, const float *__restrict__ y,
                                                         const float *__restrict__ z, uint32_t n,
                                                         const float *__restrict__ bb, uint64_t *__restrict__ keys,
                                                         uint64_t *__restrict__ sk, uint32_t *__restrict__ sv) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float bx = bb[0], by = bb[1], bz = bb[2], bs = bb[3];
    const uint64_t ix = scale_to_grid3(x[i], bx, bx + bs);
    const uint64_t iy = scale_to_grid3(y[i], by, by + bs);
    const uint64_t iz = scale_to_grid3(z[i], bz, bz + bs);
    const uint64_t k = spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
    keys[i] = k;
    sk[i] = k;
    sv[i] = i;
}

// sortParticlesByMortonKey3D's gather (morton.cpp:129-136) of every field,
// plus the (x, y, z, m) stream the walk and the tree read
__global__ __launch_bounds__(kBlock) void k3_gather(State3 src, State3 dst, uint32_t n, int sph,
                                                    const uint32_t *__restrict__ sorted_slots,
                                                    uint32_t *__restrict__ perm, float4 *__restrict__ pxyzm) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = sorted_slots[i];
    perm[i] = j;
    const float x = src.x[j], y = src.y[j], z = src.z[j], m = src.m[j];
    dst.x[i] = x; dst.y[i] = y; dst.z[i] = z; dst.m[i] = m;
    dst.vx[i] = src.vx[j]; dst.vy[i] = src.vy[j]; dst.vz[i] = src.vz[j];
    dst.ax[i] = src.ax[j]; dst.ay[i] = src.ay[j]; dst.az[i] = src.az[j];
    dst.pvx[i] = src.pvx[j]; dst.pvy[i] = src.pvy[j]; dst.pvz[i] = src.pvz[j];
    dst.flags[i] = src.flags[j];
    dst.id[i] = src.id[j];
    pxyzm[i] = make_float4(x, y, z, m);
    if (sph) {
        dst.sph_mass[i] = src.sph_mass[j]; dst.rest_dens[i] = src.rest_dens[j];
        dst.stiff[i] = src.stiff[j]; dst.visc[i] = src.visc[j];
        dst.cohesion[i] = src.cohesion[j]; dst.dens[i] = src.dens[j];
        dst.press[i] = src.press[j]; dst.press_tmp[i] = src.press_tmp[j];
        dst.pred_dens[i] = src.pred_dens[j];
        dst.ppx[i] = src.ppx[j]; dst.ppy[i] = src.ppy[j]; dst.ppz[i] = src.ppz[j];
        dst.pfx[i] = src.pfx[j]; dst.pfy[i] = src.pfy[j]; dst.pfz[i] = src.pfz[j];
    }
    if (src.color) {  // render fields
        dst.color[i] = src.color[j]; dst.pcolor[i] = src.pcolor[j]; dst.scolor[i] = src.scolor[j];
        dst.size[i] = src.size[j]; dst.nbrs[i] = src.nbrs[j];
        dst.prev_size[i] = src.prev_size[j];
    }
}

__global__ __launch_bounds__(kBlock) void k3_pack(State3 s, uint32_t n, float4 *__restrict__ pxyzm) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) pxyzm[i] = make_float4(s.x[i], s.y[i], s.z[i], s.m[i]);
}

__device__ __forceinline__ void kahan_add(float &sum, float &comp, float v) {
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// Octree accelerations: the per-particle pre-order walk of
// Physics3D::calculateForceFromGrid3D (physics3D.cpp:131-254) by a whole wave
// in lockstep, as the 2D walk: every lane follows its own reference path, the
// wave visits their union, a lane that accepted node k is muted until the
// walk leaves k's subtree, and node reads are wave-uniform scalar loads.  The
// opening test uses the reference's unfused arithmetic with the softening
// added unsquared (distanceSq = |d|^2 + softening), so the accepted-node sets
// match the CPU path; a leaf is opened by its particles with j != i.
template <int BS, bool KAHAN, bool COUNT>
__global__ __launch_bounds__(BS) void k3_bh(const float4 *__restrict__ pxyzm, const NodeRec3 *__restrict__ nodes,
                                            const uint32_t *__restrict__ counters, uint32_t n, float G,
                                            float theta_sq, float eps, State3 s, float half_dt_end, Domain3 dom,
                                            unsigned long long *__restrict__ inter) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * BS + threadIdx.x;
    const bool per = dom.w > 0.0f;
    const float hw = dom.w * 0.5f, hh = dom.h * 0.5f, hd = dom.d * 0.5f;
    const bool live = i < n;
    const float4 p = live ? pxyzm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t tree_err = counters[kCntError] & (kErrNodeCap | kErrDepth | kErrFrontier);
    const uint32_t node_count = tree_err ? 0u : __builtin_amdgcn_readfirstlane(counters[kCntNodeTotal]);
    uint32_t skip_until = 0;
    float fx = 0.f, fy = 0.f, fz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    uint32_t cnt = 0, cnt_nodes = 0;
    uint32_t idx = 0;
    const uint32_t end = live ? node_count : 0u;
    while (idx < node_count) {
        const float4 h0 = reinterpret_cast<const float4 *>(nodes)[2 * idx];
        const uint4 h1 = reinterpret_cast<const uint4 *>(nodes)[2 * idx + 1];
        const uint32_t next = h1.y;
        const uint32_t after = (next >= node_count - idx) ? node_count : idx + next + 1;
        if (h0.w <= 0.0f) {  // empty node (physics3D.cpp:147-150)
            idx = after;
            continue;
        }
        const bool active = idx >= skip_until && idx < end;
        float dx, dy, dz, dist_sq;
        bool accept;
        {
#pragma clang fp contract(off)
            dx = h0.x - p.x;
            dy = h0.y - p.y;
            dz = h0.z - p.z;
            if (per) {  // physics3D.cpp:160-164
                dx = wrap_d(dx, dom.w, hw);
                dy = wrap_d(dy, dom.h, hh);
                dz = wrap_d(dz, dom.d, hd);
            }
            dist_sq = dx * dx + dy * dy + dz * dz + eps;
            accept = __uint_as_float(h1.x) < theta_sq * dist_sq;
        }
        if (active && accept) {
            const float inv = __builtin_amdgcn_rsqf(dist_sq);
            const float f = G * h0.w * (inv * inv * inv);
            if (KAHAN) {
                kahan_add(fx, cx, dx * f);
                kahan_add(fy, cy, dy * f);
                kahan_add(fz, cz, dz * f);
            } else {
                fx = fmaf(dx, f, fx);
                fy = fmaf(dy, f, fy);
                fz = fmaf(dz, f, fz);
            }
            if (COUNT) {
                ++cnt;
                ++cnt_nodes;
            }
            skip_until = after;
        }
        const bool leaf = next == 0;
        const bool open_leaf = active && !accept && leaf;
        if (__ballot(open_leaf)) {
            // leaf direct sum over [start, end), j != i (physics3D.cpp:200-236)
            // two partners per trip, their loads issued together; each lane
            // still adds exactly its own partners, in order
            auto partner = [&](const float4 q, uint32_t j) {
                if (open_leaf && j != i) {
                    float px = q.x - p.x, py = q.y - p.y, pz = q.z - p.z;
                    if (per) {  // physics3D.cpp:206-210
                        px = wrap_d(px, dom.w, hw);
                        py = wrap_d(py, dom.h, hh);
                        pz = wrap_d(pz, dom.d, hd);
                    }
                    const float d2 = px * px + py * py + pz * pz + eps;
                    const float inv = __builtin_amdgcn_rsqf(d2);
                    const float f = G * q.w * (inv * inv * inv);
                    if (KAHAN) {
                        kahan_add(fx, cx, px * f);
                        kahan_add(fy, cy, py * f);
                        kahan_add(fz, cz, pz * f);
                    } else {
                        fx = fmaf(px, f, fx);
                        fy = fmaf(py, f, fy);
                        fz = fmaf(pz, f, fz);
                    }
                    if (COUNT) ++cnt;
                }
            };
            const uint32_t jend = min(h1.w, n);
            uint32_t j = h1.z;
            for (; j + 1 < jend; j += 2) {
                const float4 q0 = pxyzm[j], q1 = pxyzm[j + 1];
                partner(q0, j);
                partner(q1, j + 1);
            }
            if (j < jend) partner(pxyzm[j], j);
        }
        const bool descend = active && !accept && !leaf;
        idx = __ballot(descend) ? idx + 1u : after;
    }
    if (COUNT) {
        uint64_t c = cnt, cn = cnt_nodes;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            c += __shfl_xor(c, o, kWave);
            cn += __shfl_xor(cn, o, kWave);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(inter, (unsigned long long)c);
            atomicAdd(inter + 1, (unsigned long long)cn);
        }
    }
    if (!live) return;
    const float accx = KAHAN ? fx - cx : fx;
    const float accy = KAHAN ? fy - cy : fy;
    const float accz = KAHAN ? fz - cz : fz;
    s.ax[i] = accx;
    s.ay[i] = accy;
    s.az[i] = accz;
    if (half_dt_end != 0.0f && !(s.flags[i] & GE_FLAG_PINNED)) {
#pragma clang fp contract(off)
        s.vx[i] = s.vx[i] + accx * half_dt_end;
        s.vy[i] = s.vy[i] + accy * half_dt_end;
        s.vz[i] = s.vz[i] + accz * half_dt_end;
    }
}

// Physics3D::naiveGravity3D (physics3D.cpp:416-470): O(N^2) over every j
// (j == i contributes d = 0), softening added unsquared; sources staged
// through LDS one tile at a time
__global__ __launch_bounds__(kBlock) void k3_direct(const float4 *__restrict__ pxyzm, uint32_t n, float G, float eps,
                                                    State3 s, float half_dt_end, Domain3 dom) {
    const bool per = dom.w > 0.0f;
    const float hw = dom.w * 0.5f, hh = dom.h * 0.5f, hd = dom.d * 0.5f;
    __shared__ float4 tile[kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const float4 p = i < n ? pxyzm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float tx = 0.f, ty = 0.f, tz = 0.f;
    for (uint32_t base = 0; base < n; base += kBlock) {
        const uint32_t j = base + threadIdx.x;
        tile[threadIdx.x] = j < n ? pxyzm[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const uint32_t lim = min((uint32_t)kBlock, n - base);
        for (uint32_t k = 0; k < lim; ++k) {
            const float4 q = tile[k];
            float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            if (per) {  // physics3D.cpp:445-449
                dx = wrap_d(dx, dom.w, hw);
                dy = wrap_d(dy, dom.h, hh);
                dz = wrap_d(dz, dom.d, hd);
            }
            const float d2 = dx * dx + dy * dy + dz * dz + eps;
            const float inv = __builtin_amdgcn_rsqf(d2);
            const float f = G * q.w * (inv * inv * inv);
            tx = fmaf(dx, f, tx);
            ty = fmaf(dy, f, ty);
            tz = fmaf(dz, f, tz);
        }
        __syncthreads();
    }
    if (i >= n) return;
    s.ax[i] = tx;
    s.ay[i] = ty;
    s.az[i] = tz;
    if (half_dt_end != 0.0f && !(s.flags[i] & GE_FLAG_PINNED)) {
#pragma clang fp contract(off)
        s.vx[i] = s.vx[i] + tx * half_dt_end;
        s.vy[i] = s.vy[i] + ty * half_dt_end;
        s.vz[i] = s.vz[i] + tz * half_dt_end;
    }
}

__global__ __launch_bounds__(kBlock) void k3_zero_acc(State3 s, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    s.ax[i] = 0.0f;
    s.ay[i] = 0.0f;
    s.az[i] = 0.0f;
}

// Physics3D::integrateEnd3D (physics3D.cpp:1158-1168)
__global__ __launch_bounds__(kBlock) void k3_integrate_end(State3 s, uint32_t n, float half_dt) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || (s.flags[i] & GE_FLAG_PINNED)) return;
    s.vx[i] = s.vx[i] + s.ax[i] * half_dt;
    s.vy[i] = s.vy[i] + s.ay[i] * half_dt;
    s.vz[i] = s.vz[i] + s.az[i] * half_dt;
}

__global__ void k3_fill_u64(unsigned long long *__restrict__ p, unsigned long long v) { p[threadIdx.x] = v; }

constexpr int kDiagThreads = 256;

// one workgroup per target: FP64 sum over j != t of G m_j d / (|d|^2 + eps)^1.5
__global__ __launch_bounds__(kDiagThreads) void k3_direct_fp64(const float4 *__restrict__ pxyzm, uint32_t n,
                                                               const uint32_t *__restrict__ targets, double G,
                                                               double eps, double *__restrict__ out) {
    __shared__ double sh[kDiagThreads / 64][3];
    const uint32_t t = targets[blockIdx.x];
    const float4 p = pxyzm[t];
    const double xi = p.x, yi = p.y, zi = p.z;
    double v[3] = {0.0, 0.0, 0.0};
    for (uint32_t j = threadIdx.x; j < n; j += kDiagThreads) {
        if (j == t) continue;
        const float4 q = pxyzm[j];
        const double dx = (double)q.x - xi, dy = (double)q.y - yi, dz = (double)q.z - zi;
        const double r2 = dx * dx + dy * dy + dz * dz + eps;
        const double inv = 1.0 / (r2 * sqrt(r2));
        v[0] += q.w * dx * inv;
        v[1] += q.w * dy * inv;
        v[2] += q.w * dz * inv;
    }
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o, 64);
    }
    if (lane == 0)
        for (int c = 0; c < 3; ++c) sh[w][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0.0;
        for (int k = 0; k < kDiagThreads / 64; ++k) s += sh[k][threadIdx.x];
        out[3ull * blockIdx.x + threadIdx.x] = G * s;
    }
}

template <int BS, bool KAHAN, bool COUNT>
void launch3_bh_t(ge3_ctx *c, const ge3_step_params *p, float half_dt_end, Domain3 dom) {
    const uint32_t nb = blocks_for(c->n, BS);
    hipLaunchKernelGGL((k3_bh<BS, KAHAN, COUNT>), dim3(nb), dim3(BS), 0, c->stream, c->b.pxyzm, c->b.nodes,
                       c->b.counters, c->n, p->G, p->theta * p->theta, p->softening, c->b.st[c->b.cur],
                       half_dt_end, dom, c->b.inter);
}

template <int BS>
void launch3_bh_bs(ge3_ctx *c, const ge3_step_params *p, float half_dt_end, bool count, Domain3 dom) {
    const bool k = p->compensated != 0;
    if (k && count) launch3_bh_t<BS, true, true>(c, p, half_dt_end, dom);
    else if (k) launch3_bh_t<BS, true, false>(c, p, half_dt_end, dom);
    else if (count) launch3_bh_t<BS, false, true>(c, p, half_dt_end, dom);
    else launch3_bh_t<BS, false, false>(c, p, half_dt_end, dom);
}

Domain3 domain3_of(const ge3_step_params *p) {
    Domain3 d;
    if (p->periodic) {
        d.w = p->domain_w;
        d.h = p->domain_h;
        d.d = p->domain_d;
    }
    return d;
}

}  // namespace

void launch3_integrate_bbox(ge3_ctx *c, float dt, int do_integrate, Domain3 dom, int sph_clamp, float sph_max_vel) {
    const uint32_t nb = blocks_for(c->n, kBBoxChunk3);
    hipLaunchKernelGGL(k3_integrate_bbox, dim3(nb), dim3(kBlock), 0, c->stream, c->b.st[c->b.cur], c->n, dt,
                       do_integrate, dom, sph_clamp, sph_max_vel, c->b.partials);
    hipLaunchKernelGGL(k3_bbox_final, dim3(1), dim3(kBlock), 0, c->stream, c->b.partials, nb, c->b.bb);
}

void launch3_morton_keys(ge3_ctx *c, const float *bb_device) {
    const State3 &s = c->b.st[c->b.cur];
    hipLaunchKernelGGL(k3_morton_keys, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, s.z,
                       c->n, bb_device, c->b.keys, c->b.sk_a, c->b.sv_a);
}

void launch3_sort_gather(ge3_ctx *c) {
    radix_sort_kv(c->stream, c->b.hist, c->b.sk_a, c->b.sk_b, c->b.sv_a, c->b.sv_b, c->n, 0, kMortonBits3,
                  c->b.counters + kCntError);
    const int cur = c->b.cur;
    hipLaunchKernelGGL(k3_gather, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.st[cur],
                       c->b.st[1 - cur], c->n, c->sph ? 1 : 0, c->b.sv_a, c->b.perm, c->b.pxyzm);
    c->b.cur = 1 - cur;
}

void launch3_pack(ge3_ctx *c) {
    hipLaunchKernelGGL(k3_pack, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.st[c->b.cur], c->n,
                       c->b.pxyzm);
}

void launch3_zero_acc(ge3_ctx *c) {
    hipLaunchKernelGGL(k3_zero_acc, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.st[c->b.cur],
                       c->n);
}

void launch3_bh(ge3_ctx *c, const ge3_step_params *p, float half_dt_end, bool count) {
    const Domain3 dom = domain3_of(p);
    if (count) hipLaunchKernelGGL(k3_fill_u64, dim3(1), dim3(2), 0, c->stream, c->b.inter, 0ull);
    switch (p->walk_block) {
    case 64: launch3_bh_bs<64>(c, p, half_dt_end, count, dom); break;
    case 128: launch3_bh_bs<128>(c, p, half_dt_end, count, dom); break;
    default: launch3_bh_bs<256>(c, p, half_dt_end, count, dom); break;
    }
}

void launch3_direct(ge3_ctx *c, const ge3_step_params *p, float half_dt_end) {
    hipLaunchKernelGGL(k3_direct, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.pxyzm, c->n, p->G,
                       p->softening, c->b.st[c->b.cur], half_dt_end, domain3_of(p));
}

void launch3_integrate_end(ge3_ctx *c, float dt) {
    hipLaunchKernelGGL(k3_integrate_end, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream,
                       c->b.st[c->b.cur], c->n, dt * 0.5f);
}

void launch3_direct_fp64(ge3_ctx *c, double G, double eps, uint32_t nt, const uint32_t *targets, double *out) {
    hipLaunchKernelGGL(k3_direct_fp64, dim3(nt), dim3(kDiagThreads), 0, c->stream, c->b.pxyzm, c->n, targets, G, eps,
                       out);
}

}  // namespace ge
