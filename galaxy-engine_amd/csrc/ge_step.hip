// ge_step.hip -- per-frame streaming kernels of the 2D step:
//   integrateStart fused with the bounding-box reduction, Morton keys, the
//   Morton-order gather, the wave-coherent Barnes-Hut walk (with the
//   integrateEnd kick fused in), the direct sum, and integrateEnd.
#include <cfloat>

#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = glm_min(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = glm_max(v, __shfl_xor(v, o, kWave));
    return v;
}

// Block-wide min/max of four values; result valid in thread 0.
__device__ __forceinline__ void block_minmax(float &a, float &b, float &c, float &d) {
    __shared__ float red[4][kBlock / kWave];
    a = wave_min(a); b = wave_min(b); c = wave_max(c); d = wave_max(d);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = c; red[3][w] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / kWave; ++k) {
            a = glm_min(a, red[0][k]); b = glm_min(b, red[1][k]);
            c = glm_max(c, red[2][k]); d = glm_max(d, red[3][k]);
        }
    }
}

// Physics::integrateStart (physics.cpp:1069-1112) followed by the
// boundingBox() reduction (globalLogic.cpp:1083-1099) over the drifted
// positions: per-workgroup partial min/max here, the final box in
// k_bbox_final (a second launch instead of a device-scope fence).
__global__ __launch_bounds__(kBlock) void k_integrate_bbox(State s, const uint32_t *__restrict__ chunk_lo, float dt,
                                                           int sph_clamp, float max_vel, int do_integrate,
                                                           Domain dom, float *__restrict__ partials) {
#pragma clang fp contract(off)
    const float half_dt = dt * 0.5f;
    const float max_vsq = max_vel * max_vel;
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    // one workgroup per chunk; a chunk never spans two systems
    const uint32_t lo = chunk_lo[blockIdx.x], hi = chunk_lo[blockIdx.x + 1];
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        float x = s.x[i], y = s.y[i];
        if (do_integrate && !(s.flags[i] & GE_FLAG_PINNED)) {
            float pvx = s.vx[i], pvy = s.vy[i];
            float vx = pvx + s.ax[i] * half_dt;
            float vy = pvy + s.ay[i] * half_dt;
            if (sph_clamp) {
                float vsq = vx * vx + vy * vy;
                if (vsq > max_vsq) {
                    float psq = pvx * pvx + pvy * pvy;
                    if (psq > 0.00001f) {
                        float inv = max_vel / sqrtf(psq);
                        pvx = pvx * inv;
                        pvy = pvy * inv;
                    }
                    float inv = max_vel / sqrtf(vsq);
                    vx = vx * inv;
                    vy = vy * inv;
                }
            }
            s.pvx[i] = pvx; s.pvy[i] = pvy;
            s.vx[i] = vx; s.vy[i] = vy;
            if (s.hx) {
                // precision form: the drift moves the offset from the home cell
                int32_t hx = s.hx[i], hy = s.hy[i];
                float ox = s.ox[i], oy = s.oy[i];
                home_sync(x, hx, ox);
                home_sync(y, hy, oy);
                ox = ox + vx * dt;
                oy = oy + vy * dt;
                home_norm(hx, ox);
                home_norm(hy, oy);
                x = home_view(hx, ox);
                y = home_view(hy, oy);
                if (dom.w > 0.0f) {  // the wrap of physics.cpp:1104-1110, by whole domains
                    const float sx = x < 0.0f ? 1.0f : (x >= dom.w ? -1.0f : 0.0f);
                    const float sy = y < 0.0f ? 1.0f : (y >= dom.h ? -1.0f : 0.0f);
                    if (sx != 0.0f) home_shift(hx, ox, sx * dom.w);
                    if (sy != 0.0f) home_shift(hy, oy, sy * dom.h);
                    x = home_view(hx, ox);
                    y = home_view(hy, oy);
                }
                s.hx[i] = hx; s.hy[i] = hy;
                s.ox[i] = ox; s.oy[i] = oy;
            } else {
                x = x + vx * dt;
                y = y + vy * dt;
                if (dom.w > 0.0f) {  // physics.cpp:1104-1110
                    if (x < 0.0f) x += dom.w;
                    else if (x >= dom.w) x -= dom.w;
                    if (y < 0.0f) y += dom.h;
                    else if (y >= dom.h) y -= dom.h;
                }
            }
            s.x[i] = x; s.y[i] = y;
        }
        mnx = glm_min(mnx, x); mny = glm_min(mny, y);
        mxx = glm_max(mxx, x); mxy = glm_max(mxy, y);
    }
    block_minmax(mnx, mny, mxx, mxy);
    if (threadIdx.x == 0) reinterpret_cast<float4 *>(partials)[blockIdx.x] = make_float4(mnx, mny, mxx, mxy);
}

// one workgroup per system: its chunks' partials -> its box
__global__ __launch_bounds__(kBlock) void k_bbox_final(const float *__restrict__ partials,
                                                       const uint32_t *__restrict__ chunk_first,
                                                       float *__restrict__ sys_bb) {
#pragma clang fp contract(off)
    float *bb = sys_bb + 8ull * blockIdx.x;
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    const uint32_t k0 = chunk_first[blockIdx.x], k1 = chunk_first[blockIdx.x + 1];
    for (uint32_t k = k0 + threadIdx.x; k < k1; k += kBlock) {
        const float4 v = reinterpret_cast<const float4 *>(partials)[k];
        mnx = glm_min(mnx, v.x); mny = glm_min(mny, v.y);
        mxx = glm_max(mxx, v.z); mxy = glm_max(mxy, v.w);
    }
    block_minmax(mnx, mny, mxx, mxy);
    if (threadIdx.x == 0) {
        float size = glm_max(mxx - mnx, mxy - mny);
        float cx = (mnx + mxx) * 0.5f;
        float cy = (mny + mxy) * 0.5f;
        float half = size * 0.5f;
        bb[0] = cx - half;
        bb[1] = cy - half;
        bb[2] = size;
        bb[3] = 0.0f;
        // the exact extremes, for the SPH grid frame
        bb[4] = mnx; bb[5] = mny; bb[6] = mxx; bb[7] = mxy;
    }
}

// Morton::scaleToGrid / spreadBits / morton2D / computeMortonKeys
// (morton.cpp:3-38).  Writes the 36-bit key per slot and the packed sort word.
__device__ __forceinline__ uint64_t scale_to_grid(float pos, float lo, float hi) {
#pragma clang fp contract(off)
    if (hi <= lo) return 0;
    float c = (pos < lo) ? lo : ((hi < pos) ? hi : pos);
    float norm = (c - lo) / (hi - lo);
    uint64_t s = (uint64_t)(norm * 262144.0f);
    return s < 262143u ? s : 262143u;
}

__device__ __forceinline__ uint64_t spread_bits(uint64_t v) {
    v &= 0x3FFFFull;
    v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
    v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

// packed sort word: system | key | slot within the system (the system part
// is empty for a single scene), so one LSD sort orders each system by key
struct PackFmt {
    int idx_bits;
    int sys_shift;  // 0 when there is one system
};

__global__ __launch_bounds__(kBlock) void k_morton_keys(const float *__restrict__ x,
                                                        const float *__restrict__ y, uint32_t n,
                                                        const uint32_t *__restrict__ sys_of,
                                                        const uint32_t *__restrict__ sys_off,
                                                        const float *__restrict__ sys_bb, PackFmt fmt,
                                                        uint64_t *__restrict__ keys,
                                                        uint64_t *__restrict__ packed) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t sy = sys_of[i];
    const float *bb = sys_bb + 8ull * sy;
    const float bx = bb[0], by = bb[1], bs = bb[2];
    const float ext = bs < 1e-6f ? 1e-6f : bs;
    const float maxx = bx + ext, maxy = by + ext;
    uint64_t ix = scale_to_grid(x[i], bx, maxx);
    uint64_t iy = scale_to_grid(y[i], by, maxy);
    uint64_t k = spread_bits(ix) | (spread_bits(iy) << 1);
    keys[i] = k;
    const uint64_t local = i - sys_off[sy];
    packed[i] = (fmt.sys_shift ? ((uint64_t)sy << fmt.sys_shift) : 0ull) | (k << fmt.idx_bits) | local;
}

// Morton::sortParticlesByMortonKey's gather (morton.cpp:56-62) for every
// SoA field, plus the packed (x, y, m) stream the traversal reads.
__global__ __launch_bounds__(kBlock) void k_gather(State src, State dst, uint32_t n, int sph,
                                                   const uint64_t *__restrict__ sorted,
                                                   const uint32_t *__restrict__ sys_off, PackFmt fmt,
                                                   uint32_t *__restrict__ perm,
                                                   float4 *__restrict__ pxym, float4 *__restrict__ pho) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint64_t w = sorted[i];
    const uint32_t sy = fmt.sys_shift ? (uint32_t)(w >> fmt.sys_shift) : 0u;
    const uint32_t j = sys_off[sy] + (uint32_t)(w & ((1ull << fmt.idx_bits) - 1ull));
    perm[i] = j;
    float x = src.x[j], y = src.y[j], m = src.m[j];
    dst.x[i] = x; dst.y[i] = y; dst.m[i] = m;
    dst.vx[i] = src.vx[j]; dst.vy[i] = src.vy[j];
    dst.ax[i] = src.ax[j]; dst.ay[i] = src.ay[j];
    dst.pvx[i] = src.pvx[j]; dst.pvy[i] = src.pvy[j];
    dst.flags[i] = src.flags[j]; dst.id[i] = src.id[j];
    pxym[i] = make_float4(x, y, m, 0.0f);
    if (sph) {
        dst.sph_mass[i] = src.sph_mass[j]; dst.rest_dens[i] = src.rest_dens[j];
        dst.stiff[i] = src.stiff[j]; dst.visc[i] = src.visc[j];
        dst.cohesion[i] = src.cohesion[j]; dst.dens[i] = src.dens[j];
        dst.press[i] = src.press[j]; dst.press_tmp[i] = src.press_tmp[j];
        dst.pred_dens[i] = src.pred_dens[j];
        dst.ppx[i] = src.ppx[j]; dst.ppy[i] = src.ppy[j];
        dst.pfx[i] = src.pfx[j]; dst.pfy[i] = src.pfy[j];
    }
    if (src.color) {  // render fields
        dst.color[i] = src.color[j]; dst.pcolor[i] = src.pcolor[j]; dst.scolor[i] = src.scolor[j];
        dst.size[i] = src.size[j]; dst.nbrs[i] = src.nbrs[j];
        dst.prev_size[i] = src.prev_size[j];
    }
    if (src.temp) {  // temperature fields
        dst.temp[i] = src.temp[j]; dst.label[i] = src.label[j];
        dst.ke[i] = src.ke[j]; dst.prev_ke[i] = src.prev_ke[j];
    }
    if (src.hx) {  // precision form, made to agree with pos first
        int32_t hx = src.hx[j], hy = src.hy[j];
        float ox = src.ox[j], oy = src.oy[j];
        home_sync(x, hx, ox);
        home_sync(y, hy, oy);
        dst.hx[i] = hx; dst.hy[i] = hy;
        dst.ox[i] = ox; dst.oy[i] = oy;
        pho[i] = make_float4(ox, oy, __int_as_float(hx), __int_as_float(hy));
    }
}

__device__ __forceinline__ void kahan_add(float &sum, float &comp, float v) {
    float y = v - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// The walks' wave layout.  Several systems: system s owns the node range
// [P[off_s], P[off_s+1]) (P is the node-start prefix of the tree build), and
// waves are laid out per system: wave w of system s holds its particles
// 64 (w - W_s) .. 64 (w - W_s) + 63 (W the per-system wave prefix), so a wave
// holds the same particles whether the system runs alone or in a batch.
//
// !PERSIST: workgroup b of the launch takes tile xcd_block(b) once.
// PERSIST: the launch is persistent.  The layout's workgroup tiles (BS/64
// consecutive waves, the tile a one-shot launch gives a workgroup) are cut
// into kWalkQueues contiguous ranges; workgroup b starts on range
// b mod kWalkQueues (the XCD that hardware dispatch gives it, so
// neighbouring tiles share an L2) and, when its range runs dry, takes tiles
// from the next ranges.  walk_q[0..kWalkQueues) are the ranges' tickets and
// walk_q[kWalkQueues] counts the waves that have finished; the last wave out
// clears them all for the next launch (walk_retire).
template <bool PERSIST>
__device__ __forceinline__ uint32_t walk_next_tile(uint32_t total_tiles, uint32_t *walk_q, uint32_t &q_at,
                                                   uint32_t &q_seen, bool &first, uint32_t *tile_s) {
    if (!PERSIST) {
        if (!first) return kNone;
        first = false;
        return xcd_block(blockIdx.x, gridDim.x);
    }
    while (q_seen < kWalkQueues) {
        const uint32_t lo = (uint32_t)((uint64_t)total_tiles * q_at / kWalkQueues);
        const uint32_t hi = (uint32_t)((uint64_t)total_tiles * (q_at + 1) / kWalkQueues);
        __syncthreads();  // every wave has read the previous ticket
        if (threadIdx.x == 0) *tile_s = atomicAdd(&walk_q[q_at], 1u);
        __syncthreads();
        const uint32_t t = *tile_s;
        if (t < hi - lo) return lo + t;
        q_at = (q_at + 1) % kWalkQueues;
        ++q_seen;
    }
    return kNone;
}

template <int BS, bool PERSIST>
__device__ __forceinline__ void walk_retire(uint32_t *walk_q) {
    if (PERSIST && (threadIdx.x & (kWave - 1)) == 0) {
        const uint32_t waves = gridDim.x * (BS / kWave);
        if (atomicAdd(&walk_q[kWalkQueues], 1u) == waves - 1u)
            for (uint32_t k = 0; k <= kWalkQueues; ++k) walk_q[k] = 0u;
    }
}

// the system of wave w: the last s with W_s <= w (wave-uniform search)
__device__ __forceinline__ uint32_t wave_system(const uint32_t *__restrict__ sys_wave_off, uint32_t nsys,
                                                uint32_t w) {
    uint32_t slo = 0, shi = nsys;
    while (shi - slo > 1u) {
        const uint32_t mid = (slo + shi) >> 1;
        if (sys_wave_off[mid] <= w) slo = mid;
        else shi = mid;
    }
    return slo;
}

// Barnes-Hut accelerations: the per-particle pre-order walk of
// Physics::calculateForceFromGrid (physics.cpp:91-205), executed by a whole
// wave in lockstep.  Every lane walks its own reference path; the wave visits
// the union of the 64 paths.  A lane that accepted node k stays muted until
// the walk leaves k's subtree (index k + next + 1).  Node reads are
// wave-uniform (scalar loads).  The acceptance test uses the reference's
// unfused arithmetic so the accepted-node sets match the CPU path.  TEMP adds
// the walk's node-level heat transfer (physics.cpp:261-276).
template <int BS, bool KAHAN, bool COUNT, bool TEMP, bool PERSIST, bool PREC>
__global__ __launch_bounds__(BS, 8) void k_bh(const float4 *__restrict__ pxym,
                                           const NodeRec *__restrict__ nodes,
                                           const uint32_t *__restrict__ start_prefix,
                                           const uint32_t *__restrict__ sys_wave_off, uint32_t nsys,
                                           const uint32_t *__restrict__ sys_off,
                                           const uint32_t *__restrict__ counters, uint32_t n,
                                           float G, float theta_sq, float soft_sq,
                                           float *__restrict__ ax, float *__restrict__ ay,
                                           float *__restrict__ vx, float *__restrict__ vy,
                                           const uint32_t *__restrict__ flags, float half_dt_end,
                                           Domain dom, unsigned long long *__restrict__ inter,
                                           const float *__restrict__ gtavg, float *__restrict__ temp,
                                           float heat_k, float heat_dt, uint32_t *__restrict__ walk_q,
                                           const float4 *__restrict__ pho) {
    __shared__ uint32_t tile_s;
    const uint32_t total_tiles = (sys_wave_off[nsys] + (BS / kWave) - 1) / (BS / kWave);
    uint32_t q_at = blockIdx.x % kWalkQueues, q_seen = 0;
    bool first = true;
    for (;;) {
        const uint32_t blk = walk_next_tile<PERSIST>(total_tiles, walk_q, q_at, q_seen, first, &tile_s);
        if (blk == kNone) break;
        const uint32_t w = __builtin_amdgcn_readfirstlane(blk * (BS / kWave) + (threadIdx.x / kWave));
        const uint32_t sy = wave_system(sys_wave_off, nsys, w);
        const uint32_t i = sys_off[sy] + (w - sys_wave_off[sy]) * kWave + (threadIdx.x & (kWave - 1));
        const bool per = dom.w > 0.0f;
        const float hw = dom.w * 0.5f, hh = dom.h * 0.5f;
        const bool live = w < sys_wave_off[nsys] && i < sys_off[sy + 1];
        const float4 p = live ? pxym[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        // PREC: the home form (ox, oy, hx, hy) of the target; separations come
        // from home forms (the precision form)
        const float4 ph = (PREC && live) ? pho[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int32_t phx = __float_as_int(ph.z), phy = __float_as_int(ph.w);
        float ti = (TEMP && live) ? temp[i] : 0.0f;  // temp[i], moved by every accepted node
        // a tree that raised an error flag may hold unwritten nodes: do not walk it
        const uint32_t tree_err = counters[kCntError] & (kErrNodeCap | kErrDepth | kErrFrontier);
        const uint32_t node_count = tree_err ? 0u : __builtin_amdgcn_readfirstlane(counters[kCntNodeTotal]);
        uint32_t root = kNone, end = 0u;
        if (live) {
            root = start_prefix[sys_off[sy]];
            end = min(start_prefix[sys_off[sy + 1]], node_count);
        }
        uint32_t wlo = root, whi = end;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            wlo = min(wlo, (uint32_t)__shfl_xor((int)wlo, o, kWave));
            whi = max(whi, (uint32_t)__shfl_xor((int)whi, o, kWave));
        }
        const uint32_t walk_end = __builtin_amdgcn_readfirstlane(whi);
        uint32_t skip_until = root;
        float fx = 0.f, fy = 0.f, cx = 0.f, cy = 0.f;
        uint32_t cnt = 0, cnt_nodes = 0, visits = 0, leaf_trips = 0;  // visits and trips: per wave
        uint32_t idx = __builtin_amdgcn_readfirstlane(wlo);
        while (idx < walk_end) {
            const NodeRec nd = nodes[idx];
            const uint32_t next = nd.next;
            if (COUNT) ++visits;
            // index after this subtree, saturated so the walk always moves forward
            const uint32_t after = (next >= node_count - idx) ? node_count : idx + next + 1;
            if (nd.mass <= 0.0f) {  // empty node: skip the subtree (physics.cpp:107-110)
                idx = after;
                continue;
            }
            const bool active = idx >= skip_until && idx < end;
            float dx, dy, dist_sq;
            bool accept;
            {
#pragma clang fp contract(off)
                dx = nd.comx - p.x;
                dy = nd.comy - p.y;
                if (per) {  // scalar variant (physics.cpp:117-120)
                    dx = wrap_d(dx, dom.w, hw);
                    dy = wrap_d(dy, dom.h, hh);
                }
                dist_sq = dx * dx + dy * dy + soft_sq;
                accept = nd.size_sq < theta_sq * dist_sq;
            }
            const bool leaf = next == 0;
            const bool open_leaf = active && !accept && leaf;
            if (active && accept) {
                const float inv = __builtin_amdgcn_rsqf(dist_sq);
                const float inv2 = inv * inv;
                const float inv3 = inv2 * inv;
                const float f = G * nd.mass * inv3;
                if (KAHAN) {
                    kahan_add(fx, cx, dx * f);
                    kahan_add(fy, cy, dy * f);
                } else {
                    fx += dx * f;
                    fy += dy * f;
                }
                if (COUNT) {
                    ++cnt;
                    ++cnt_nodes;
                }
                if (TEMP) {  // heat transfer toward the node's average (physics.cpp:261-276)
#pragma clang fp contract(off)
                    const uint32_t cn = nd.end - nd.start;
                    if (cn > 0u) {
                        const float avg = gtavg[idx];  // gridTemp / cn, divided once per node
                        const float diff = avg - ti;
                        const float invd = 1.0f / sqrtf(dist_sq);
                        float dist = 0.0f;
                        if (dist_sq > 1e-16f) dist = 1.0f / invd;
                        if (dist > 1e-8f) {
                            const float heat = heat_k * diff / dist;
                            ti = ti + heat * heat_dt;
                        }
                    }
                }
                skip_until = after;
            }
            if (__ballot(open_leaf)) {
                // leaf direct sum over [start, end), skipping i == j (physics.cpp:163-196)
                const uint32_t jend = min(nd.end, n);
                if (COUNT) leaf_trips += jend > nd.start ? jend - nd.start : 0u;
                // two partners per trip, their loads issued together; each lane
                // still adds exactly its own partners, in order
                auto partner = [&](const float4 q, const float4 qh, uint32_t j) {
                    if (open_leaf && j != i) {
                        float px, py;
                        if (PREC) {
                            px = home_sep(__float_as_int(qh.z), qh.x, phx, ph.x);
                            py = home_sep(__float_as_int(qh.w), qh.y, phy, ph.y);
                        } else {
                            px = q.x - p.x;
                            py = q.y - p.y;
                        }
                        if (per) {  // physics.cpp:169-172
                            px = wrap_d(px, dom.w, hw);
                            py = wrap_d(py, dom.h, hh);
                        }
                        const float d2 = px * px + py * py + soft_sq;
                        const float inv = __builtin_amdgcn_rsqf(d2);
                        const float f = G * q.z * (inv * inv * inv);
                        if (KAHAN) {
                            kahan_add(fx, cx, px * f);
                            kahan_add(fy, cy, py * f);
                        } else {
                            fx = fmaf(px, f, fx);
                            fy = fmaf(py, f, fy);
                        }
                        if (COUNT) ++cnt;
                    }
                };
                const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
                uint32_t j = nd.start;
                for (; j + 1 < jend; j += 2) {
                    const float4 q0 = pxym[j], q1 = pxym[j + 1];
                    const float4 qh0 = PREC ? pho[j] : zero4, qh1 = PREC ? pho[j + 1] : zero4;
                    partner(q0, qh0, j);
                    partner(q1, qh1, j + 1);
                }
                if (j < jend) partner(pxym[j], PREC ? pho[j] : zero4, j);
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
                atomicAdd(inter + 2, (unsigned long long)visits);
                atomicAdd(inter + 3, (unsigned long long)leaf_trips);
            }
        }
        if (live) {
            if (TEMP) temp[i] = ti;
            const float accx = KAHAN ? fx - cx : fx;
            const float accy = KAHAN ? fy - cy : fy;
            ax[i] = accx;
            ay[i] = accy;
            if (half_dt_end != 0.0f && !(flags[i] & GE_FLAG_PINNED)) {
#pragma clang fp contract(off)
                vx[i] = vx[i] + accx * half_dt_end;
                vy[i] = vy[i] + accy * half_dt_end;
            }
        }
    }
    walk_retire<BS, PERSIST>(walk_q);
}

// Acceleration of one accepted node in the accuracy modes (ORDER 2 / 3 / 4,
// no counterpart in the reference): the Taylor expansion of the softened
// kernel about the node's FP64 centre of mass up to the quadrupole (2),
// octupole (3) or hexadecapole (4) term.  With d = target - COM,
// R^2 = |d|^2 + eps^2, r = 1/R, k = r^2, A_n = C_n / n! and the traces T, U,
// W of A3 and A4:
//   a = G r^3 (S d - k V)
//   S = -M + k (3 tr A2 + k (F2 + k (F3 + k F4)))
//   V = g2 + k (g3 + k g4)
//   g2 = -6 A2 d + 9 T,  g3 = -45 A3:dd + 180 U d,  g4 = -420 A4:ddd
//   F2 = -15 A2:dd + 45 T.d - 45 W = 2.5 d.(g2 + 9 T) - 45 W
//   F3 = -105 A3:ddd + 630 U:dd = (7/3) (d.g3 + d.(180 U d) / 2)
//   F4 = -945 A4:dddd = (9/4) d.g4
// (the Cartesian Taylor series, D_n = (-1)^n (2n-1)!! r^(2n+1); the scalar
// parts F_n follow from the gradients g_n by Euler's rule for homogeneous
// polynomials).  The node record holds the coefficients of g2, g3 and g4
// and the constants (multipole.hip's finish_record), so the evaluation is
// three short polynomials and two Horner chains in k.  qx, qy = d; inv = r.
template <int ORDER>
__device__ __forceinline__ void mp_accel(float4 h0, float4 h1, float4 h2, float4 h3, float4 h4, float4 h5, float qx,
                                         float qy, float inv, float G, float &tx, float &ty) {
    const float k = inv * inv;
    const float A = G * (inv * k);
    float S, vx, vy;
    if (ORDER == 2) {
        const float gx = fmaf(h1.x, qx, h1.y * qy), gy = fmaf(h1.y, qx, h1.z * qy);
        const float f2 = 2.5f * fmaf(qx, gx, qy * gy);
        S = fmaf(k, fmaf(k, f2, h1.w), -h0.w);
        vx = gx;
        vy = gy;
    } else {
        const float xx = qx * qx, yy = qy * qy, xy2 = (qx + qx) * qy;
        // g2 (with 9 T) and F2
        const float g2x = fmaf(h1.x, qx, fmaf(h1.y, qy, h2.x));
        const float g2y = fmaf(h1.y, qx, fmaf(h1.z, qy, h2.y));
        const float d2t = fmaf(qx, g2x + h2.x, qy * (g2y + h2.y));
        if (ORDER == 3) {  // the series without the fourth-order moments (A4, U, W)
            const float g3x = fmaf(h3.x, xx, fmaf(h3.y, xy2, h3.z * yy));
            const float g3y = fmaf(h3.y, xx, fmaf(h3.z, xy2, h3.w * yy));
            const float f3 = (7.0f / 3.0f) * fmaf(qx, g3x, qy * g3y);
            S = fmaf(k, fmaf(k, fmaf(k, f3, 2.5f * d2t), h1.w), -h0.w);
            vx = fmaf(k, g3x, g2x);
            vy = fmaf(k, g3y, g2y);
        } else {
            const float f2 = fmaf(2.5f, d2t, h2.z);
            // g3 and the U part of F3
            const float ux = fmaf(h4.x, qx, h2.w * qy), uy = fmaf(h2.w, qx, h4.y * qy);
            const float g3x = fmaf(h3.x, xx, fmaf(h3.y, xy2, fmaf(h3.z, yy, ux)));
            const float g3y = fmaf(h3.y, xx, fmaf(h3.z, xy2, fmaf(h3.w, yy, uy)));
            const float t3 = fmaf(qx, fmaf(0.5f, ux, g3x), qy * fmaf(0.5f, uy, g3y));  // (3/7) F3
            // g4 (h5.x and h5.w carry the binomial weight 3; h5.y's is applied here)
            const float x3 = xx * qx, y3 = yy * qy, a = xx * qy, b = qx * yy, b3 = 3.0f * b;
            const float g4x = fmaf(h4.z, x3, fmaf(h5.w, a, fmaf(h5.x, b, h5.y * y3)));
            const float g4y = fmaf(h4.w, x3, fmaf(h5.x, a, fmaf(h5.y, b3, h5.z * y3)));
            const float dg4 = fmaf(qx, g4x, qy * g4y);                                    // (4/9) F4
            const float in3 = fmaf(k * (27.0f / 28.0f), dg4, t3);                          // (3/7) (F3 + k F4)
            const float in2 = fmaf(k * (7.0f / 3.0f), in3, f2);                            // F2 + k (F3 + k F4)
            S = fmaf(k, fmaf(k, in2, h1.w), -h0.w);
            vx = fmaf(k, fmaf(k, g4x, g3x), g2x);
            vy = fmaf(k, fmaf(k, g4y, g3y), g2y);
        }
    }
    tx = A * fmaf(qx, S, -k * vx);
    ty = A * fmaf(qy, S, -k * vy);
}

// Leaf direct sum of the accuracy modes' walk: sx, sy += m_j d_j r^-3 over
// particles [j0, j1), two pairs per trip with their loads issued together.
// With softening the self pair adds exactly zero (its separation is zero),
// so SELF (the self test) is only needed without softening.
template <bool SELF, bool PREC>
__device__ __forceinline__ void leaf_sum(const float4 *__restrict__ pxym, const float4 *__restrict__ pho, uint32_t j0,
                                         uint32_t j1, float4 p, float4 ph, uint32_t self, bool per, Domain dom,
                                         float hw, float hh, float soft_sq, float &sx, float &sy) {
    const int32_t phx = __float_as_int(ph.z), phy = __float_as_int(ph.w);
    auto pair = [&](const float4 q, const float4 qh, uint32_t j) {
        float px, py;
        if (PREC) {
            px = home_sep(__float_as_int(qh.z), qh.x, phx, ph.x);
            py = home_sep(__float_as_int(qh.w), qh.y, phy, ph.y);
        } else {
            px = q.x - p.x;
            py = q.y - p.y;
        }
        if (per) {
            px = wrap_d(px, dom.w, hw);
            py = wrap_d(py, dom.h, hh);
        }
        const float inv = __builtin_amdgcn_rsqf(fmaf(px, px, fmaf(py, py, soft_sq)));
        float f = q.z * (inv * inv * inv);
        if (SELF) f = j != self ? f : 0.0f;
        sx = fmaf(px, f, sx);
        sy = fmaf(py, f, sy);
    };
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t j = j0;
    for (; j + 1 < j1; j += 2) {
        const float4 q0 = pxym[j], q1 = pxym[j + 1];
        const float4 qh0 = PREC ? pho[j] : zero4, qh1 = PREC ? pho[j + 1] : zero4;
        pair(q0, qh0, j);
        pair(q1, qh1, j + 1);
    }
    if (j < j1) pair(pxym[j], PREC ? pho[j] : zero4, j);
}

// The walk of the accuracy modes (no counterpart in the reference).  A node
// is accepted when |d| exceeds its opening radius (size/theta + COM offset,
// or b_max/theta) for every lane of the wave, and then every lane adds its
// expansion (mp_accel); when some lane fails, the wave descends, and a leaf
// that any lane opens is summed directly by every lane.  Each lane therefore
// sums over a partition of its system into nodes whose opening test it passes
// and single particles, a refinement of the partition its own walk would
// take; every block runs with all lanes busy, and the walk keeps no per-lane
// state.  Lanes past the end of the system walk as copies of its last
// particle and write nothing.  The node record and its moments are read with
// scalar loads.  With softening the self pair adds exactly zero (its
// separation is zero), so the leaf loop needs no self test; without
// softening it keeps one.  KAHAN: the accepted-node terms between two leaves
// are summed plainly and added, with the leaf's sum, to a compensated sum.
template <int BS, bool KAHAN, bool COUNT, int ORDER, bool PERSIST, bool PREC>
__global__ __launch_bounds__(BS, 8) void k_bh_mp(const float4 *__restrict__ pxym,
                                              const NodeRec *__restrict__ nodes,
                                              const float4 *__restrict__ mp,
                                              const uint32_t *__restrict__ start_prefix,
                                              const uint32_t *__restrict__ sys_wave_off, uint32_t nsys,
                                              const uint32_t *__restrict__ sys_off,
                                              const uint32_t *__restrict__ counters, uint32_t n, float G,
                                              float soft_sq, float *__restrict__ ax, float *__restrict__ ay,
                                              float *__restrict__ vx, float *__restrict__ vy,
                                              const uint32_t *__restrict__ flags, float half_dt_end, Domain dom,
                                              unsigned long long *__restrict__ inter,
                                              uint32_t *__restrict__ walk_q, const float4 *__restrict__ pho,
                                              const float4 *__restrict__ node_home) {
    __shared__ uint32_t tile_s;
    const uint32_t total_waves = sys_wave_off[nsys];
    const uint32_t total_tiles = (total_waves + (BS / kWave) - 1) / (BS / kWave);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t q_at = blockIdx.x % kWalkQueues, q_seen = 0;
    bool first = true;
    for (;;) {
        const uint32_t blk = walk_next_tile<PERSIST>(total_tiles, walk_q, q_at, q_seen, first, &tile_s);
        if (blk == kNone) break;
        const uint32_t w = __builtin_amdgcn_readfirstlane(blk * (BS / kWave) + (threadIdx.x / kWave));
        if (w >= total_waves) continue;
        const uint32_t sy = wave_system(sys_wave_off, nsys, w);
        const uint32_t s_lo = sys_off[sy], s_hi = sys_off[sy + 1];
        const uint32_t i = s_lo + (w - sys_wave_off[sy]) * kWave + lane;
        const bool live = i < s_hi;
        const uint32_t ic = live ? i : s_hi - 1u;  // the wave's first lane is always live
        const bool per = dom.w > 0.0f;
        const float hw = dom.w * 0.5f, hh = dom.h * 0.5f;
        const float4 p = pxym[ic];
        const float4 ph = PREC ? pho[ic] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int32_t phx = __float_as_int(ph.z), phy = __float_as_int(ph.w);
        // a tree that raised an error flag may hold unwritten nodes: do not walk it
        const uint32_t tree_err = counters[kCntError] & (kErrNodeCap | kErrDepth | kErrFrontier | kErrMultipole);
        const uint32_t node_count = tree_err ? 0u : __builtin_amdgcn_readfirstlane(counters[kCntNodeTotal]);
        const uint32_t root = __builtin_amdgcn_readfirstlane(start_prefix[s_lo]);
        const uint32_t end = __builtin_amdgcn_readfirstlane(min(start_prefix[s_hi], node_count));
        float fx = 0.f, fy = 0.f, cx = 0.f, cy = 0.f;
        float nfx = 0.f, nfy = 0.f;  // KAHAN: accepted-node terms since the last leaf
        uint32_t cnt = 0, cnt_nodes = 0, visits = 0, leaf_trips = 0;  // visits and trips: per wave
        uint32_t idx = root;
        while (idx < end) {
            // the node, its whole multipole record and (PREC) its centre's
            // home form are loaded together, before any branch
            const NodeRec nd = nodes[idx];
            const float4 *rec = mp + 6ull * idx;
            const float4 h0 = rec[0], h1 = rec[1];
            float4 h2 = make_float4(0.f, 0.f, 0.f, 0.f), h3 = h2, h4 = h2, h5 = h2;
            if (ORDER >= 3) {
                h2 = rec[2];
                h3 = rec[3];
            }
            if (ORDER >= 4) {
                h4 = rec[4];
                h5 = rec[5];
                // keeps these loads in the same round trip as the node's
                // instead of a second one inside the accepting branch
                asm volatile("" ::"s"(h4.x), "s"(h4.y), "s"(h4.z), "s"(h4.w), "s"(h5.x), "s"(h5.y), "s"(h5.z),
                             "s"(h5.w));
            }
            float4 hc;  // PREC: the expansion centre's home form
            if (PREC) hc = node_home[idx];
            const uint32_t next = nd.next;
            if (COUNT) ++visits;
            // index after this subtree, saturated so the walk always moves forward
            const uint32_t after = (next >= node_count - idx) ? node_count : idx + next + 1;
            // empty node (mass <= 0, NaN excluded), tested on the scalar unit:
            // its subtree is skipped
            const uint32_t mb = __float_as_uint(nd.mass);
            const bool empty = (int32_t)mb <= 0 && (mb & 0x7fffffffu) <= 0x7f800000u;
            float dx, dy;  // COM - target
            if (PREC) {
                dx = home_sep(__float_as_int(hc.z), hc.x, phx, ph.x);
                dy = home_sep(__float_as_int(hc.w), hc.y, phy, ph.y);
            } else {
                dx = h0.x - p.x;
                dy = h0.y - p.y;
            }
            if (per) {
                dx = wrap_d(dx, dom.w, hw);
                dy = wrap_d(dy, dom.h, hh);
            }
            const float d2 = dx * dx + dy * dy;
            const bool take = !empty && !__ballot(!(d2 > h0.z));  // h0.z: the squared opening radius
            const bool open = !empty && !take && next == 0;
            if (take) {
                float tx, ty;
                mp_accel<ORDER>(h0, h1, h2, h3, h4, h5, -dx, -dy, __builtin_amdgcn_rsqf(d2 + soft_sq), G, tx, ty);
                if (KAHAN) {  // plain within a node run, compensated per run
                    nfx += tx;
                    nfy += ty;
                } else {
                    fx += tx;
                    fy += ty;
                }
                if (COUNT) {
                    cnt += live ? 1u : 0u;
                    cnt_nodes += live ? 1u : 0u;
                }
            }
            if (open) {
                // leaf direct sum: the pairs go to a plain per-leaf sum of
                // m_j d_j r^-3, scaled by G and compensated once per leaf
                const uint32_t jend = min(nd.end, n);
                if (COUNT) {  // pairs j != i of live lanes
                    const uint32_t len = jend > nd.start ? jend - nd.start : 0u;
                    leaf_trips += len;
                    if (live) cnt += len - ((nd.start <= i && i < jend) ? 1u : 0u);
                }
                float sx = 0.f, sy = 0.f;
                if (soft_sq > 0.0f)
                    leaf_sum<false, PREC>(pxym, pho, nd.start, jend, p, ph, ic, per, dom, hw, hh, soft_sq, sx, sy);
                else
                    leaf_sum<true, PREC>(pxym, pho, nd.start, jend, p, ph, ic, per, dom, hw, hh, soft_sq, sx, sy);
                if (KAHAN) {  // the node run since the last leaf and this leaf's sum
                    kahan_add(fx, cx, fmaf(G, sx, nfx));
                    kahan_add(fy, cy, fmaf(G, sy, nfy));
                    nfx = 0.f;
                    nfy = 0.f;
                } else {
                    fx = fmaf(G, sx, fx);
                    fy = fmaf(G, sy, fy);
                }
            }
            idx = (empty || take || open) ? after : idx + 1u;
        }
        if (COUNT) {
            uint64_t c = cnt, cn = cnt_nodes;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                c += __shfl_xor(c, o, kWave);
                cn += __shfl_xor(cn, o, kWave);
            }
            if (lane == 0) {
                atomicAdd(inter, (unsigned long long)c);
                atomicAdd(inter + 1, (unsigned long long)cn);
                atomicAdd(inter + 2, (unsigned long long)visits);
                atomicAdd(inter + 3, (unsigned long long)leaf_trips);
            }
        }
        if (live) {
            if (KAHAN) {
                kahan_add(fx, cx, nfx);
                kahan_add(fy, cy, nfy);
            }
            const float accx = KAHAN ? fx - cx : fx;
            const float accy = KAHAN ? fy - cy : fy;
            ax[i] = accx;
            ay[i] = accy;
            if (half_dt_end != 0.0f && !(flags[i] & GE_FLAG_PINNED)) {
#pragma clang fp contract(off)
                vx[i] = vx[i] + accx * half_dt_end;
                vy[i] = vy[i] + accy * half_dt_end;
            }
        }
    }
    walk_retire<BS, PERSIST>(walk_q);
}

// Physics::naiveGravity (physics.cpp:392-434): O(N^2) with the softening
// added unsquared; sources staged through LDS one tile at a time.  PREC: the
// separations come from the home forms (the precision form).
template <bool PREC>
__global__ __launch_bounds__(kBlock) void k_direct(const float4 *__restrict__ pxym, const float4 *__restrict__ pho,
                                                   uint32_t n, float G, float eps, float *__restrict__ ax,
                                                   float *__restrict__ ay, float *__restrict__ vx,
                                                   float *__restrict__ vy,
                                                   const uint32_t *__restrict__ flags,
                                                   float half_dt_end, Domain dom) {
    const bool per = dom.w > 0.0f;
    const float hw = dom.w * 0.5f, hh = dom.h * 0.5f;
    __shared__ float4 tile[kBlock];
    __shared__ float4 htile[PREC ? kBlock : 1];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const float4 p = i < n ? pxym[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 ph = (PREC && i < n) ? pho[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float tx = 0.f, ty = 0.f;
    for (uint32_t base = 0; base < n; base += kBlock) {
        const uint32_t j = base + threadIdx.x;
        tile[threadIdx.x] = j < n ? pxym[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (PREC) htile[threadIdx.x] = j < n ? pho[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const uint32_t lim = min((uint32_t)kBlock, n - base);
        for (uint32_t k = 0; k < lim; ++k) {
            const float4 q = tile[k];
            float dx, dy;
            if (PREC) {
                const float4 qh = htile[k];
                dx = home_sep(__float_as_int(qh.z), qh.x, __float_as_int(ph.z), ph.x);
                dy = home_sep(__float_as_int(qh.w), qh.y, __float_as_int(ph.w), ph.y);
            } else {
                dx = q.x - p.x;
                dy = q.y - p.y;
            }
            if (per) {  // physics.cpp:417-420
                dx = wrap_d(dx, dom.w, hw);
                dy = wrap_d(dy, dom.h, hh);
            }
            const float d2 = dx * dx + dy * dy + eps;
            const float inv = __builtin_amdgcn_rsqf(d2);
            const float f = G * q.z * (inv * inv * inv);
            tx = fmaf(dx, f, tx);
            ty = fmaf(dy, f, ty);
        }
        __syncthreads();
    }
    if (i >= n) return;
    ax[i] = tx;
    ay[i] = ty;
    if (half_dt_end != 0.0f && !(flags[i] & GE_FLAG_PINNED)) {
#pragma clang fp contract(off)
        vx[i] = vx[i] + tx * half_dt_end;
        vy[i] = vy[i] + ty * half_dt_end;
    }
}

__global__ void k_fill_u64(uint64_t *__restrict__ p, uint64_t v) { p[threadIdx.x] = v; }

// sys_of[i] = the system holding slot i (binary search over the offsets)
__global__ __launch_bounds__(kBlock) void k_sys_of(const uint32_t *__restrict__ sys_off, uint32_t nsys, uint32_t n,
                                                   uint32_t *__restrict__ sys_of) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = nsys;  // sys_off[lo] <= i < sys_off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sys_off[mid] <= i) lo = mid;
        else hi = mid;
    }
    sys_of[i] = lo;
}

// Moves kept systems to their new slots: new slot i of system t takes old
// slot src_off[t] + (i - sys_off[t]).  Every field moves, so the state is the
// same as if the system had never shared the context.
__global__ __launch_bounds__(kBlock) void k_compact(State src, State dst, uint32_t n, int sph,
                                                    const uint32_t *__restrict__ sys_of,
                                                    const uint32_t *__restrict__ sys_off,
                                                    const uint32_t *__restrict__ src_off) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = sys_of[i];
    const uint32_t j = src_off[t] + (i - sys_off[t]);
    dst.x[i] = src.x[j]; dst.y[i] = src.y[j]; dst.m[i] = src.m[j];
    dst.vx[i] = src.vx[j]; dst.vy[i] = src.vy[j];
    dst.ax[i] = src.ax[j]; dst.ay[i] = src.ay[j];
    dst.pvx[i] = src.pvx[j]; dst.pvy[i] = src.pvy[j];
    dst.flags[i] = src.flags[j]; dst.id[i] = src.id[j];
    if (sph) {
        dst.sph_mass[i] = src.sph_mass[j]; dst.rest_dens[i] = src.rest_dens[j];
        dst.stiff[i] = src.stiff[j]; dst.visc[i] = src.visc[j];
        dst.cohesion[i] = src.cohesion[j]; dst.dens[i] = src.dens[j];
        dst.press[i] = src.press[j]; dst.press_tmp[i] = src.press_tmp[j];
        dst.pred_dens[i] = src.pred_dens[j];
        dst.ppx[i] = src.ppx[j]; dst.ppy[i] = src.ppy[j];
        dst.pfx[i] = src.pfx[j]; dst.pfy[i] = src.pfy[j];
    }
    if (src.color) {  // render fields
        dst.color[i] = src.color[j]; dst.pcolor[i] = src.pcolor[j]; dst.scolor[i] = src.scolor[j];
        dst.size[i] = src.size[j]; dst.nbrs[i] = src.nbrs[j];
        dst.prev_size[i] = src.prev_size[j];
    }
    if (src.temp) {  // temperature fields
        dst.temp[i] = src.temp[j]; dst.label[i] = src.label[j];
        dst.ke[i] = src.ke[j]; dst.prev_ke[i] = src.prev_ke[j];
    }
    if (src.hx) {  // precision form
        dst.hx[i] = src.hx[j]; dst.hy[i] = src.hy[j];
        dst.ox[i] = src.ox[j]; dst.oy[i] = src.oy[j];
    }
}

__global__ __launch_bounds__(kBlock) void k_zero2(float *__restrict__ a, float *__restrict__ b, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) { a[i] = 0.f; b[i] = 0.f; }
}

// SPH::groundModeBoundary (SPH.cpp:283-322): vertical gravity and four
// walls at distance radiusMultiplier (= h) from the domain edges
__global__ __launch_bounds__(kBlock) void k_sph_ground(State s, uint32_t n, float radius, float w, float h,
                                                       float vertical_gravity, float friction) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || (s.flags[i] & GE_FLAG_PINNED)) return;
    const float damp = -0.1f;  // SPH::boundDamping (SPH.h:18)
    float x = s.x[i], y = s.y[i], vx = s.vx[i], vy = s.vy[i];
    s.ay[i] = s.ay[i] + vertical_gravity;
    if (x - radius < 0.0f) { vx *= damp; vy *= 1.0f - friction; x = radius; }
    if (x + radius > w) { vx *= damp; vy *= 1.0f - friction; x = w - radius; }
    if (y - radius < 0.0f) { vy *= damp; vx *= 1.0f - friction; y = radius; }
    if (y + radius > h) { vy *= damp; vx *= 1.0f - friction; y = h - radius; }
    s.x[i] = x; s.y[i] = y; s.vx[i] = vx; s.vy[i] = vy;
}

// Physics::integrateEnd (physics.cpp:1114-1125)
__global__ __launch_bounds__(kBlock) void k_integrate_end(State s, uint32_t n, float half_dt) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || (s.flags[i] & GE_FLAG_PINNED)) return;
    s.vx[i] = s.vx[i] + s.ax[i] * half_dt;
    s.vy[i] = s.vy[i] + s.ay[i] * half_dt;
}

}  // namespace

Domain domain_of(const ge_step_params *p) {
    Domain d;
    if (p->periodic) {
        d.w = p->domain_w;
        d.h = p->domain_h;
    }
    return d;
}

void launch_integrate_start_bbox(ge_ctx *c, float dt, int sph_clamp, float max_vel, int do_integrate, Domain dom) {
    hipLaunchKernelGGL(k_integrate_bbox, dim3(c->nchunks), dim3(kBlock), 0, c->stream, c->b.st[c->b.cur],
                       c->b.chunk_lo, dt, sph_clamp, max_vel, do_integrate, dom, c->b.partials);
    hipLaunchKernelGGL(k_bbox_final, dim3(c->nsys), dim3(kBlock), 0, c->stream, c->b.partials, c->b.chunk_first,
                       c->b.sys_bb);
}

static PackFmt pack_fmt(const ge_ctx *c) {
    return PackFmt{c->idx_bits, c->sys_bits ? c->idx_bits + kMortonBits : 0};
}

void launch_morton_keys(ge_ctx *c) {
    const State &s = c->b.st[c->b.cur];
    hipLaunchKernelGGL(k_morton_keys, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, c->n,
                       c->b.sys_of, c->b.sys_off, c->b.sys_bb, pack_fmt(c), c->b.keys, c->b.sort_a);
}

void launch_sort_keys(ge_ctx *c) {
    radix_sort_packed(c, c->b.sort_a, c->b.sort_b, c->n, c->idx_bits, c->idx_bits + kMortonBits + c->sys_bits);
}

void launch_gather(ge_ctx *c) {
    const int cur = c->b.cur;
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.st[cur],
                       c->b.st[1 - cur], c->n, (int)c->sph, c->b.sort_a, c->b.sys_off, pack_fmt(c), c->b.perm,
                       c->b.pxym, c->b.pho);
    c->b.cur = 1 - cur;
}

template <int BS, bool KAHAN, bool COUNT, int ORDER, bool TEMP, bool PERSIST, bool PREC>
void launch_bh_p(ge_ctx *c, const ge_step_params *p, float end_kick_half_dt, uint32_t nb, uint32_t *walk_q) {
    const State &s = c->b.st[c->b.cur];
    const float theta_sq = p->theta * p->theta;
    const float soft_sq = p->softening * p->softening;
    auto *inter = reinterpret_cast<unsigned long long *>(c->b.interactions);
    if constexpr (ORDER == 0)
        hipLaunchKernelGGL((k_bh<BS, KAHAN, COUNT, TEMP, PERSIST, PREC>), dim3(nb), dim3(BS), 0, c->stream, c->b.pxym,
                           c->b.nodes, c->b.start_count, c->b.sys_wave_off, c->nsys, c->b.sys_off, c->b.counters,
                           c->n, p->G, theta_sq, soft_sq, s.ax, s.ay, s.vx, s.vy, s.flags, end_kick_half_dt,
                           domain_of(p), inter, c->b.node_tavg, s.temp, c->thermal_p.heat_conductivity, p->dt, walk_q,
                           c->b.pho);
    else
        hipLaunchKernelGGL((k_bh_mp<BS, KAHAN, COUNT, ORDER, PERSIST, PREC>), dim3(nb), dim3(BS), 0, c->stream,
                           c->b.pxym, c->b.nodes, c->b.node_mp, c->b.start_count, c->b.sys_wave_off, c->nsys,
                           c->b.sys_off, c->b.counters, c->n, p->G, soft_sq, s.ax, s.ay, s.vx, s.vy, s.flags,
                           end_kick_half_dt, domain_of(p), inter, walk_q, c->b.pho, c->b.node_home);
}

template <int BS, bool KAHAN, bool COUNT, int ORDER, bool TEMP, bool PERSIST>
void launch_bh_k(ge_ctx *c, const ge_step_params *p, float end_kick_half_dt, uint32_t nb, uint32_t *walk_q) {
    if (c->precise) launch_bh_p<BS, KAHAN, COUNT, ORDER, TEMP, PERSIST, true>(c, p, end_kick_half_dt, nb, walk_q);
    else launch_bh_p<BS, KAHAN, COUNT, ORDER, TEMP, PERSIST, false>(c, p, end_kick_half_dt, nb, walk_q);
}

// the persistent form of a walk kernel, for the occupancy query
template <int BS, bool KAHAN, bool COUNT, int ORDER, bool TEMP>
const void *persistent_walk() {
    if constexpr (ORDER == 0) return reinterpret_cast<const void *>(k_bh<BS, KAHAN, COUNT, TEMP, true, false>);
    else return reinterpret_cast<const void *>(k_bh_mp<BS, KAHAN, COUNT, ORDER, true, false>);
}

template <int BS, bool KAHAN, bool COUNT, int ORDER, bool TEMP>
void launch_bh_t(ge_ctx *c, const ge_step_params *p, float end_kick_half_dt) {
    const uint32_t nb = blocks_for((uint64_t)c->nwaves * kWave, BS);
    if (c->nsys <= 1) {
        launch_bh_k<BS, KAHAN, COUNT, ORDER, TEMP, false>(c, p, end_kick_half_dt, nb, nullptr);
        return;
    }
    // a batch of systems runs one persistent launch: as many workgroups as
    // the device holds at once, each pulling waves from the ticket ranges
    static int per_cu = 0;  // one value per kernel instance
    if (per_cu == 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, persistent_walk<BS, KAHAN, COUNT, ORDER, TEMP>(), BS, 0) !=
                hipSuccess || v <= 0)
            v = 1;
        per_cu = v;
    }
    const uint64_t resident = (uint64_t)per_cu * c->cus;
    launch_bh_k<BS, KAHAN, COUNT, ORDER, TEMP, true>(c, p, end_kick_half_dt, resident < nb ? (uint32_t)resident : nb,
                                                     c->b.counters + kCntWalkQ0);
}

template <int BS, int ORDER, bool TEMP>
void launch_bh_o(ge_ctx *c, const ge_step_params *p, float end_kick_half_dt, bool count) {
    const bool k = p->compensated != 0;
    if (k && count) launch_bh_t<BS, true, true, ORDER, TEMP>(c, p, end_kick_half_dt);
    else if (k) launch_bh_t<BS, true, false, ORDER, TEMP>(c, p, end_kick_half_dt);
    else if (count) launch_bh_t<BS, false, true, ORDER, TEMP>(c, p, end_kick_half_dt);
    else launch_bh_t<BS, false, false, ORDER, TEMP>(c, p, end_kick_half_dt);
}

template <int BS>
void launch_bh_bs(ge_ctx *c, const ge_step_params *p, float end_kick_half_dt, bool count) {
    if (p->accuracy == 2) launch_bh_o<BS, 4, false>(c, p, end_kick_half_dt, count);
    else if (p->accuracy == 3) launch_bh_o<BS, 3, false>(c, p, end_kick_half_dt, count);
    else if (p->accuracy == 1) launch_bh_o<BS, 2, false>(c, p, end_kick_half_dt, count);
    else if (p->temperature) launch_bh_o<BS, 0, true>(c, p, end_kick_half_dt, count);  // the reference walk only
    else launch_bh_o<BS, 0, false>(c, p, end_kick_half_dt, count);
}

void launch_bh(ge_ctx *c, const ge_step_params *p, float end_kick_half_dt, bool count) {
    if (count) hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(4), 0, c->stream, c->b.interactions, 0ull);
    switch (p->walk_block) {
    case 64: launch_bh_bs<64>(c, p, end_kick_half_dt, count); break;
    case 128: launch_bh_bs<128>(c, p, end_kick_half_dt, count); break;
    default: launch_bh_bs<256>(c, p, end_kick_half_dt, count); break;
    }
}

void launch_direct(ge_ctx *c, const ge_step_params *p, float end_kick_half_dt) {
    const State &s = c->b.st[c->b.cur];
    if (c->precise)
        hipLaunchKernelGGL(k_direct<true>, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.pxym,
                           c->b.pho, c->n, p->G, p->softening, s.ax, s.ay, s.vx, s.vy, s.flags, end_kick_half_dt,
                           domain_of(p));
    else
        hipLaunchKernelGGL(k_direct<false>, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.pxym,
                           c->b.pho, c->n, p->G, p->softening, s.ax, s.ay, s.vx, s.vy, s.flags, end_kick_half_dt,
                           domain_of(p));
}

void launch_sys_of(ge_ctx *c) {
    hipLaunchKernelGGL(k_sys_of, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.sys_off, c->nsys,
                       c->n, c->b.sys_of);
}

void launch_compact(ge_ctx *c, uint32_t n_kept) {
    const int cur = c->b.cur;
    hipLaunchKernelGGL(k_compact, dim3(blocks_for(n_kept, kBlock)), dim3(kBlock), 0, c->stream, c->b.st[cur],
                       c->b.st[1 - cur], n_kept, (int)c->sph, c->b.sys_of, c->b.sys_off, c->b.sys_src);
    c->b.cur = 1 - cur;
}

void launch_zero_acc(ge_ctx *c) {
    const State &s = c->b.st[c->b.cur];
    hipLaunchKernelGGL(k_zero2, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s.ax, s.ay, c->n);
}

void launch_sph_ground(ge_ctx *c, const ge_step_params *p) {
    hipLaunchKernelGGL(k_sph_ground, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->b.st[c->b.cur],
                       c->n, p->sph_h, p->domain_w, p->domain_h, p->vertical_gravity, p->boundary_friction);
}

void launch_integrate_end(ge_ctx *c, float dt) {
    hipLaunchKernelGGL(k_integrate_end, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream,
                       c->b.st[c->b.cur], c->n, dt * 0.5f);
}

}  // namespace ge
