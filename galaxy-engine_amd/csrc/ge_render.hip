// ge_render.hip -- the steps either side of the physics frame on the device
// (include/ge_render.h): the render fields' upload and download, the
// post-step consumers of the neighbour counts and accelerations
// (DensitySize::sizeByDensity, GE/include/Particles/densitySize.h:14-79;
// the solid, density and force modes of ColorVisuals::particlesColorVisuals,
// GE/include/Particles/particleColorVisuals.h:111-277), the vertex
// compaction of the PLY export (GE/src/globalLogic.cpp:146-242) and the
// galaxy spawner (GE/src/Particles/particlesSpawning.cpp:130-244).
//
// Every consumer is one pass over the particles: HBM-bound streaming (a
// few dozen bytes per particle), no reuse to stage.  The PLY text itself is
// formatted on the host from the compacted vertices (20 bytes per visible
// particle come back instead of the whole state), with the same printf
// conversions the reference uses, in parallel chunks.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ge_internal3d.hpp"

namespace ge {
namespace {

constexpr float kPiF = 3.14159265358979323846f;  // raylib PI, a float
constexpr uint32_t kWhite = 0xFFFFFFFFu;

__global__ __launch_bounds__(kBlock) void k_render_fill(uint32_t *__restrict__ color, uint32_t *__restrict__ pcolor,
                                                        uint32_t *__restrict__ scolor, float *__restrict__ size,
                                                        float *__restrict__ prev_size, int32_t *__restrict__ nbrs,
                                                        uint32_t at, uint32_t cnt, uint32_t rgba, float sz) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cnt) return;
    const uint32_t i = at + k;
    color[i] = rgba;
    pcolor[i] = kWhite;
    scolor[i] = kWhite;
    size[i] = sz;
    prev_size[i] = sz;  // previousSize = size (particle.h:121)
    nbrs[i] = 0;
}

// ---- consumers -----------------------------------------------------------

struct VisConst {
    int density_size, force_size, solid_color, density_color, force_color;
    float min_size, max_size, size_acc, size_mult;
    int size_max_nb;
    uint32_t p_color, s_color;
    float inv_color_max_nb;  // 1.0f / maxNeighbors (particleColorVisuals.h:146)
    float min_color_acc, max_color_acc;
};

// std::clamp and std::min as the reference calls them
__device__ __forceinline__ float std_clamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
__device__ __forceinline__ float std_min(float a, float b) { return b < a ? b : a; }

// raymath Lerp: start + amount * (end - start)
__device__ __forceinline__ float rl_lerp(float a, float b, float t) {
#pragma clang fp contract(off)
    return a + t * (b - a);
}

// raylib ColorLerp: the factor clamped to [0, 1], each channel
// (unsigned char)((1 - f) * c1 + f * c2)
__device__ __forceinline__ uint32_t rl_color_lerp(uint32_t c1, uint32_t c2, float f) {
#pragma clang fp contract(off)
    if (f < 0.0f) f = 0.0f;
    else if (f > 1.0f) f = 1.0f;
    uint32_t out = 0;
    for (int ch = 0; ch < 4; ++ch) {
        const float a = (float)((c1 >> (8 * ch)) & 255u), b = (float)((c2 >> (8 * ch)) & 255u);
        const float v = (1.0f - f) * a + f * b;
        out |= ((uint32_t)v & 255u) << (8 * ch);
    }
    return out;
}

// One pass: sizeByDensity (force size, then density size), then the colour
// modes in particlesColorVisuals' order (solid, density, force).  az is NULL
// in 2D.
__global__ __launch_bounds__(kBlock) void k_visuals(uint32_t n, const uint32_t *__restrict__ flags,
                                                    const float *__restrict__ ax, const float *__restrict__ ay,
                                                    const float *__restrict__ az, uint32_t *__restrict__ color,
                                                    uint32_t *__restrict__ pcolor, uint32_t *__restrict__ scolor,
                                                    float *__restrict__ size, const int32_t *__restrict__ nbrs,
                                                    VisConst V) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = flags[i];
    const bool dark = (f & GE_FLAG_DARK) != 0, solid = (f & GE_FLAG_SOLID) != 0;
    const bool unique = (f & GE_FLAG_UNIQUE_COLOR) != 0;
    float accSq = 0.0f;
    if (V.force_size || V.force_color) {
        const float x = ax[i], y = ay[i];
        accSq = x * x + y * y;
        if (az) {
            const float z = az[i];
            accSq = accSq + z * z;
        }
    }
    const int32_t nb = nbrs[i];
    // DensitySize::sizeByDensity (densitySize.h:16-79)
    if (V.force_size && !(solid || dark)) {
        const float clampedAcc = std_clamp(sqrtf(accSq), 0.0f, V.size_acc);
        const float norm = clampedAcc / V.size_acc;
        size[i] = rl_lerp(V.max_size * V.size_mult, V.min_size * V.size_mult, norm);
    }
    if (V.density_size && !(dark || solid)) {
        const float nd = std_min((float)nb / (float)V.size_max_nb, 1.0f);
        // (float)pow(normalDensity, 2): the double square of a float, rounded once
        size[i] = rl_lerp(V.max_size * V.size_mult, V.min_size * V.size_mult, nd * nd);
    }
    // ColorVisuals::particlesColorVisuals (particleColorVisuals.h:111-277)
    uint32_t pc = pcolor[i], sc = scolor[i];
    if (V.solid_color) {
        if (!unique) pc = V.p_color;
        color[i] = pc;
    }
    if (V.density_color && !dark) {
        if (!unique) {
            pc = V.p_color;
            sc = V.s_color;
        }
        const float nd = std_min((float)nb * V.inv_color_max_nb, 1.0f);
        color[i] = rl_color_lerp(pc, sc, nd);
    }
    if (V.force_color && !dark) {
        const float clampedAcc = std_clamp(sqrtf(accSq), V.min_color_acc, V.max_color_acc);
        const float norm = clampedAcc / V.max_color_acc;
        if (!unique) {
            pc = V.p_color;
            sc = V.s_color;
        }
        color[i] = rl_color_lerp(pc, sc, norm);
    }
    if (V.solid_color || V.density_color || V.force_color) {
        pcolor[i] = pc;
        scolor[i] = sc;
    }
}

VisConst make_vis(const ge_visual_params *v) {
    VisConst V;
    V.density_size = v->density_size != 0;
    V.force_size = v->force_size != 0;
    V.solid_color = v->solid_color != 0;
    V.density_color = v->density_color != 0;
    V.force_color = v->force_color != 0;
    V.min_size = v->min_size;
    V.max_size = v->max_size;
    V.size_acc = v->size_acc;
    V.size_mult = v->size_multiplier;
    V.size_max_nb = v->size_max_neighbors;
    V.p_color = v->p_color;
    V.s_color = v->s_color;
    V.inv_color_max_nb = 1.0f / (float)v->color_max_neighbors;
    V.min_color_acc = v->min_color_acc;
    V.max_color_acc = v->max_color_acc;
    return V;
}

// ---- PLY vertices: order-preserving compaction of the visible particles ---

constexpr int kPlyBlock = 256;

__global__ __launch_bounds__(kPlyBlock) void k_ply_count(uint32_t n, const uint32_t *__restrict__ flags,
                                                         uint32_t *__restrict__ block_cnt) {
    __shared__ uint32_t wsum[kPlyBlock / 64];
    const uint32_t i = blockIdx.x * kPlyBlock + threadIdx.x;
    const bool keep = i < n && !(flags[i] & GE_FLAG_DARK);
    const uint64_t m = __ballot(keep);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kPlyBlock / 64; ++w) t += wsum[w];
        block_cnt[blockIdx.x] = t;
    }
}

// exclusive scan of nb block counts in place, total in cnt[nb]; one workgroup
__global__ __launch_bounds__(1024) void k_ply_scan(uint32_t nb, uint32_t *__restrict__ cnt) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < nb ? cnt[j] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
            const uint32_t add = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (j < nb) cnt[j] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[nb] = carry;
}

// plyFileCreation (globalLogic.cpp:172-181) in 2D (az == NULL: scaled into
// the domain, z = 0) or plyFileCreation3D (globalLogic.cpp:222-231)
__global__ __launch_bounds__(kPlyBlock) void k_ply_write(uint32_t n, const uint32_t *__restrict__ flags,
                                                         const float *__restrict__ x, const float *__restrict__ y,
                                                         const float *__restrict__ z, const uint32_t *__restrict__ color,
                                                         const float *__restrict__ size,
                                                         const uint32_t *__restrict__ block_off, float dw, float dh,
                                                         ge_ply_vertex *__restrict__ out) {
#pragma clang fp contract(off)
    __shared__ uint32_t wsum[kPlyBlock / 64];
    const uint32_t i = blockIdx.x * kPlyBlock + threadIdx.x;
    const bool keep = i < n && !(flags[i] & GE_FLAG_DARK);
    const uint64_t m = __ballot(keep);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (!keep) return;
    uint32_t o = block_off[blockIdx.x];
    for (uint32_t k = 0; k < w; ++k) o += wsum[k];
    o += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    ge_ply_vertex v;
    if (z) {
        v.x = x[i];
        v.y = y[i];
        v.z = z[i];
    } else {
        float px = ((x[i] / dw) * 2.0f) - 1.0f;
        float py = ((y[i] / dh) * 2.0f) - 1.0f;
        const float ratio = dw / dh;
        px = px * ratio;
        px = px * 10.0f;
        py = py * 10.0f;
        v.x = px;
        v.y = py;
        v.z = 0.0f;
    }
    const uint32_t c = color[i];
    v.r = (uint8_t)(c & 255u);
    v.g = (uint8_t)((c >> 8) & 255u);
    v.b = (uint8_t)((c >> 16) & 255u);
    v.pad = 0;
    v.radius = size[i];
    out[o] = v;
}

// device vertices of the current particles; returns the count
int ply_collect(hipStream_t st, uint32_t n, const uint32_t *flags, const float *x, const float *y, const float *z,
                const uint32_t *color, const float *size, float dw, float dh, std::vector<ge_ply_vertex> *host,
                uint32_t cap, ge_ply_vertex *out, uint32_t *count, std::string &err) {
    const uint32_t nb = (uint32_t)((n + kPlyBlock - 1) / kPlyBlock);
    uint32_t *off = nullptr;
    ge_ply_vertex *dv = nullptr;
    auto bail = [&](const char *m) {
        if (off) (void)hipFree(off);
        if (dv) (void)hipFree(dv);
        err = m;
        return GE_E_HIP;
    };
    if (hipMalloc(&off, sizeof(uint32_t) * ((size_t)nb + 1)) != hipSuccess) return bail("vertex scan allocation failed");
    hipLaunchKernelGGL(k_ply_count, dim3(nb), dim3(kPlyBlock), 0, st, n, flags, off);
    hipLaunchKernelGGL(k_ply_scan, dim3(1), dim3(1024), 0, st, nb, off);
    uint32_t total = 0;
    if (hipMemcpyAsync(&total, off + nb, sizeof total, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return bail("vertex count copy failed");
    *count = total;
    const bool want = host || (out && cap > 0);
    if (want && total > 0) {
        if (hipMalloc(&dv, sizeof(ge_ply_vertex) * (size_t)total) != hipSuccess)
            return bail("vertex buffer allocation failed");
        hipLaunchKernelGGL(k_ply_write, dim3(nb), dim3(kPlyBlock), 0, st, n, flags, x, y, z, color, size, off, dw, dh,
                           dv);
        ge_ply_vertex *dst = out;
        uint32_t m = std::min(total, cap);
        if (host) {
            host->resize(total);
            dst = host->data();
            m = total;
        }
        if (hipMemcpyAsync(dst, dv, sizeof(ge_ply_vertex) * (size_t)m, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return bail("vertex copy failed");
    }
    if (off) (void)hipFree(off);
    if (dv) (void)hipFree(dv);
    return GE_OK;
}

// the PLY text of plyFileCreation: header and one printf line per vertex,
// formatted in parallel chunks that are then joined in order
std::string ply_text(const ge_ply_vertex *v, uint32_t count) {
    std::string head = "ply\nformat ascii 1.0\nelement vertex ";
    head += std::to_string(count);
    head += "\nproperty float x\nproperty float y\nproperty float z\nproperty uchar red\nproperty uchar green\n"
            "property uchar blue\nproperty float radius\nend_header\n";
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (count < 65536) nt = 1;
    std::vector<std::string> part(nt);
    auto work = [&](unsigned t) {
        const uint64_t a = (uint64_t)count * t / nt, b = (uint64_t)count * (t + 1) / nt;
        std::string &s = part[t];
        s.reserve((size_t)(b - a) * 48);
        char line[160];
        for (uint64_t k = a; k < b; ++k) {
            const ge_ply_vertex &q = v[k];
            const int len = std::snprintf(line, sizeof line, "%.6g %.6g %.6g %d %d %d %.6g\n", (double)q.x,
                                          (double)q.y, (double)q.z, (int)q.r, (int)q.g, (int)q.b, (double)q.radius);
            s.append(line, (size_t)len);
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, t);
        for (auto &t : th) t.join();
    }
    size_t total = head.size();
    for (auto &s : part) total += s.size();
    head.reserve(total);
    for (auto &s : part) head += s;
    return head;
}

int write_file(const char *path, const std::string &text, std::string &err) {
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        err = "cannot open the PLY file for writing";
        return GE_E_INVALID;
    }
    const size_t w = std::fwrite(text.data(), 1, text.size(), f);
    const int rc = std::fclose(f);
    if (w != text.size() || rc != 0) {
        err = "PLY write failed";
        return GE_E_INVALID;
    }
    return GE_OK;
}

// ---- galaxy spawner --------------------------------------------------------

// splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): one 64-bit output per counter value
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float draw(uint64_t seed, uint32_t stream, uint32_t k, uint32_t d,
                                      const float *__restrict__ u, uint32_t per) {
    if (u) return u[(size_t)k * per + d];
    const uint64_t z = splitmix64(seed + ((uint64_t)stream << 60) + 8ull * k + d);
    return (float)(z >> 40) * 0x1p-24f;
}

struct GalaxyConst {
    float cx, cy, evx, evy;
    uint32_t visible, dark;
    float scale_length, max_cum, rc;
    double dm_base_d;  // 1 + (R / rc)^2 in double
    float mass_visible, mass_dark, scatter;
    uint32_t first_id;
    uint64_t seed;
};

// ParticlesSpawning::particlesInitialConditions, the galaxy tool
// (particlesSpawning.cpp:136-243): visible particle k < visible, then dark
// particle k - visible.
__global__ __launch_bounds__(kBlock) void k_spawn_galaxy(State s, uint32_t at, GalaxyConst G,
                                                         const float *__restrict__ uv, const float *__restrict__ ud) {
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= G.visible + G.dark) return;
    const uint32_t i = at + k;
    float x, y, vx, vy, m;
    uint32_t fl;
    if (k < G.visible) {
        const float u0 = draw(G.seed, 0, k, 0, uv, 3), u1 = draw(G.seed, 0, k, 1, uv, 3);
        const float u2 = draw(G.seed, 0, k, 2, uv, 3);
        const float randVal = u0 * G.max_cum;
        float r = -G.scale_length * logf(1.0f - randVal);
        r = r < 0.01f ? 0.01f : r;  // std::max(finalRadius, 0.01f)
        const float angle = u1 * 2.0f * kPiF;
        const float dx = cosf(angle), dy = sinf(angle);
        x = G.cx + dx * r;
        y = G.cy + dy * r;
        const float speed = 10.0f * sqrtf(1758.0f / (r + 54.0f));
        vx = ((-dy) * speed) * 0.85f + G.evx;
        vy = (dx * speed) * 0.85f + G.evy;
        const float mult = 1.0f + (u2 * 2.0f - 1.0f) * G.scatter;
        m = G.mass_visible * mult;
        fl = GE_FLAG_SPH;  // ParticleRendering(..., isSPH = true, ...)
    } else {
        const uint32_t q = k - G.visible;
        const float u0 = draw(G.seed, 1, q, 0, ud, 5), u1 = draw(G.seed, 1, q, 1, ud, 5);
        const float u2 = draw(G.seed, 1, q, 2, ud, 5), u3 = draw(G.seed, 1, q, 3, ud, 5);
        const float u4 = draw(G.seed, 1, q, 4, ud, 5);
        // rc * sqrt((float)(pow(1 + (R / rc)^2, u) - 1)), the pow in double
        const float inner = (float)(pow(G.dm_base_d, (double)u0) - 1.0);
        const float r = G.rc * sqrtf(inner);
        const float angle = u1 * 2.0f * kPiF;
        x = G.cx + r * cosf(angle);
        y = G.cy + r * sinf(angle);
        // rand() % 60 - 30 from a draw in [0, 1)
        int ix = (int)(u2 * 60.0f), iy = (int)(u3 * 60.0f);
        ix = ix > 59 ? 59 : ix;
        iy = iy > 59 ? 59 : iy;
        vx = (float)(ix - 30) * 0.85f + G.evx;
        vy = (float)(iy - 30) * 0.85f + G.evy;
        const float mult = 1.0f + (u4 * 2.0f - 1.0f) * G.scatter;
        m = G.mass_dark * mult;
        fl = GE_FLAG_DARK | GE_FLAG_UNIQUE_COLOR;  // ParticleRendering(..., uniqueColor, ..., isDarkMatter, ...)
    }
    s.x[i] = x;
    s.y[i] = y;
    s.vx[i] = vx;
    s.vy[i] = vy;
    s.ax[i] = 0.0f;
    s.ay[i] = 0.0f;
    s.pvx[i] = 0.0f;
    s.pvy[i] = 0.0f;
    s.m[i] = m;
    s.flags[i] = fl;
    s.id[i] = G.first_id + k;
    if (s.sph_mass) {  // ParticlePhysics(pos, vel, mass, 0.008, 1, 1, 1) (particle.h:47-73)
        s.sph_mass[i] = m / 8500000000.0f;
        s.rest_dens[i] = 0.008f;
        s.stiff[i] = 1.0f;
        s.visc[i] = 1.0f;
        s.cohesion[i] = 1.0f;
        s.dens[i] = 0.0f;
        s.press[i] = 0.0f;
        s.press_tmp[i] = 0.0f;
        s.pred_dens[i] = 0.0f;
        s.ppx[i] = 0.0f;
        s.ppy[i] = 0.0f;
        s.pfx[i] = 0.0f;
        s.pfy[i] = 0.0f;
    }
    if (s.color) {  // colour {128, 128, 128, 100} (visible) or alpha 0 (dark), size 0.125
        s.color[i] = k < G.visible ? 0x64808080u : 0x00808080u;
        s.pcolor[i] = kWhite;
        s.scolor[i] = kWhite;
        s.size[i] = 0.125f;
        s.prev_size[i] = 0.125f;
        s.nbrs[i] = 0;
    }
}

int fail(std::string &err, int code, const char *msg) {
    err = msg;
    return code;
}

// the error flags the grid kernels raise (a frame too wide for the cell table)
int grid_errors(hipStream_t st, uint32_t *counters, std::string &err) {
    uint32_t e = 0;
    if (hipMemcpyAsync(&e, counters + kCntError, sizeof e, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(err, GE_E_HIP, "error flag copy failed");
    if (e) {
        const uint32_t z = 0;
        (void)hipMemcpy(counters + kCntError, &z, sizeof z, hipMemcpyHostToDevice);
        return fail(err, (e & kErrGridRange) ? GE_E_RANGE : GE_E_DEVICE,
                    (e & kErrGridRange) ? "neighbour grid: coordinates span more cells than the table holds"
                                        : "device error flags raised");
    }
    return GE_OK;
}

int check_visuals(const ge_visual_params *v, std::string &err, bool is3d) {
    if (v->count_neighbors) {
        if (v->search != 0 && v->search != 1) return fail(err, GE_E_INVALID, "search must be 0 or 1");
        if (is3d && v->search != 0)
            return fail(err, GE_E_UNSUPPORTED, "3D neighbour counts follow NeighborSearch3DV2AVX2 (search 0)");
        if (!(v->cell > 0.0f)) return fail(err, GE_E_INVALID, "cell must be > 0");
        if (v->search == 1 && !(v->radius > 0.0f && v->radius <= 8.0f * v->cell))
            return fail(err, GE_E_INVALID, "radius must be in (0, 8 cells]");
    }
    if (v->density_size && v->size_max_neighbors <= 0) return fail(err, GE_E_INVALID, "size_max_neighbors must be > 0");
    if (v->density_color && v->color_max_neighbors <= 0)
        return fail(err, GE_E_INVALID, "color_max_neighbors must be > 0");
    return GE_OK;
}

template <class Ctx>
int render_copy(Ctx *c, const ge_render_fields *f, bool up, uint32_t *color, uint32_t *pcolor, uint32_t *scolor,
                float *size, int32_t *nbrs, float *prev_size) {
    const size_t n = c->n;
    if (n == 0) return GE_OK;
    auto mv = [&](void *dev, void *host, size_t bytes) -> int {
        if (!host) return GE_OK;
        const hipError_t e = up ? hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream)
                                : hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) return fail(c->err, GE_E_HIP, "render field copy failed");
        return GE_OK;
    };
    int r = 0;
    if ((r = mv(color, f->color, 4 * n)) || (r = mv(pcolor, f->pcolor, 4 * n)) || (r = mv(scolor, f->scolor, 4 * n)) ||
        (r = mv(size, f->size, 4 * n)) || (r = mv(nbrs, f->neighbors, 4 * n)) ||
        (r = mv(prev_size, f->previous_size, 4 * n)))
        return r;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c->err, GE_E_HIP, "render field copy failed");
    return GE_OK;
}

}  // namespace

void launch_render_fill(hipStream_t st, uint32_t *color, uint32_t *pcolor, uint32_t *scolor, float *size,
                        float *prev_size, int32_t *nbrs, uint32_t at, uint32_t cnt, uint32_t rgba, float sz) {
    if (cnt == 0) return;
    hipLaunchKernelGGL(k_render_fill, dim3(blocks_for(cnt, kBlock)), dim3(kBlock), 0, st, color, pcolor, scolor, size,
                       prev_size, nbrs, at, cnt, rgba, sz);
}

void launch_spawn_galaxy(ge_ctx *c, const State &s, uint32_t at, const ge_galaxy_params *g, uint64_t seed,
                         const float *d_uniforms) {
    GalaxyConst G;
    G.cx = g->center_x;
    G.cy = g->center_y;
    G.evx = g->extra_vx;
    G.evy = g->extra_vy;
    G.visible = g->visible;
    G.dark = g->dark;
    G.scale_length = g->scale_length;
    const float maxRadius = g->outer_radius + 600.0f;
    G.max_cum = 1.0f - std::exp(-maxRadius / g->scale_length);  // float std::exp (particlesSpawning.cpp:139)
    G.rc = g->radius_core_dm;
    const float ratio = g->outer_radius_dm / g->radius_core_dm;
    G.dm_base_d = 1.0 + (double)ratio * (double)ratio;  // 1 + pow(R / rc, 2), in double
    G.mass_visible = g->mass_visible;
    G.mass_dark = g->mass_dark;
    G.scatter = g->mass_scatter;
    G.first_id = g->first_id;
    G.seed = seed;
    const uint32_t total = g->visible + g->dark;
    const float *uv = d_uniforms;
    const float *ud = d_uniforms ? d_uniforms + 3ull * g->visible : nullptr;
    hipLaunchKernelGGL(k_spawn_galaxy, dim3(blocks_for(total, kBlock)), dim3(kBlock), 0, c->stream, s, at, G, uv, ud);
}

}  // namespace ge

using namespace ge;

extern "C" {

void ge_default_visual_params(ge_visual_params *v) {
    if (!v) return;
    std::memset(v, 0, sizeof *v);
    v->count_neighbors = 1;
    v->search = 0;             // the x86-64 build takes the AVX2 search (globalLogic.cpp:1238-1246)
    v->cell = 3.0f;            // neighborSearch.h:604, 981
    v->radius = 3.5f;          // neighborSearch.h:441
    v->density_size = 0;       // isDensitySizeEnabled
    v->force_size = 0;
    v->min_size = 0.17f;       // densitySize.h:7-12
    v->max_size = 0.62f;
    v->size_acc = 22.0f;
    v->size_max_neighbors = 200;
    v->size_multiplier = 1.0f;
    v->solid_color = 0;        // particleColorVisuals.h:8-30
    v->density_color = 1;
    v->force_color = 0;
    v->p_color = 0x6444'2800u;  // {0, 40, 68, 100}
    v->s_color = 0x4B28'509Bu;  // {155, 80, 40, 75}
    v->color_max_neighbors = 200;
    v->min_color_acc = 0.0f;
    v->max_color_acc = 40.0f;
}

int ge_render_enabled(const ge_ctx *c, int32_t *on) {
    if (!c || !on) return GE_E_INVALID;
    *on = c->render ? 1 : 0;
    return GE_OK;
}

int ge3_render_enabled(const ge3_ctx *c, int32_t *on) {
    if (!c || !on) return GE_E_INVALID;
    *on = c->render ? 1 : 0;
    return GE_OK;
}

int ge_render_upload(ge_ctx *c, const ge_render_fields *f) {
    if (!c || !f) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    (void)hipSetDevice(c->device);
    const State &s = c->b.st[c->b.cur];
    return render_copy(c, f, true, s.color, s.pcolor, s.scolor, s.size, s.nbrs, s.prev_size);
}

int ge_render_download(ge_ctx *c, const ge_render_fields *f) {
    if (!c || !f) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    (void)hipSetDevice(c->device);
    const State &s = c->b.st[c->b.cur];
    return render_copy(c, f, false, s.color, s.pcolor, s.scolor, s.size, s.nbrs, s.prev_size);
}

int ge3_render_upload(ge3_ctx *c, const ge_render_fields *f) {
    if (!c || !f) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    (void)hipSetDevice(c->device);
    const State3 &s = c->b.st[c->b.cur];
    return render_copy(c, f, true, s.color, s.pcolor, s.scolor, s.size, s.nbrs, s.prev_size);
}

int ge3_render_download(ge3_ctx *c, const ge_render_fields *f) {
    if (!c || !f) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    (void)hipSetDevice(c->device);
    const State3 &s = c->b.st[c->b.cur];
    return render_copy(c, f, false, s.color, s.pcolor, s.scolor, s.size, s.nbrs, s.prev_size);
}

int ge_visuals(ge_ctx *c, const ge_visual_params *v) {
    if (!c || !v) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    if (int r = check_visuals(v, c->err, false)) return r;
    if (v->count_neighbors && !c->sph)
        return fail(c->err, GE_E_INVALID, "neighbour counts use the SPH grid: create the context with enable_sph");
    if (v->count_neighbors && c->nsys > 1) return fail(c->err, GE_E_UNSUPPORTED, "neighbour counts need one system");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    const State &s = c->b.st[c->b.cur];
    if (v->count_neighbors) {
        launch_integrate_start_bbox(c, 0.0f, 0, 0.0f, 0);  // the grid frame uses the current extremes
        if (int r = sph_count_into(c, v->search, v->cell, v->radius, s.nbrs)) return r;
        if (int r = grid_errors(c->stream, c->b.counters, c->err)) return r;
    }
    hipLaunchKernelGGL(k_visuals, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->n, s.flags, s.ax,
                       s.ay, nullptr, s.color, s.pcolor, s.scolor, s.size, s.nbrs, make_vis(v));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c->err, GE_E_HIP, "visuals failed");
    return GE_OK;
}

int ge3_visuals(ge3_ctx *c, const ge_visual_params *v) {
    if (!c || !v) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    if (int r = check_visuals(v, c->err, true)) return r;
    if (v->count_neighbors && !c->sph)
        return fail(c->err, GE_E_INVALID, "neighbour counts use the SPH3D grid: create the context with enable_sph");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    const State3 &s = c->b.st[c->b.cur];
    if (v->count_neighbors) sph3_count_into(c, v->cell, s.nbrs);
    hipLaunchKernelGGL(k_visuals, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->n, s.flags, s.ax,
                       s.ay, s.az, s.color, s.pcolor, s.scolor, s.size, s.nbrs, make_vis(v));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c->err, GE_E_HIP, "visuals failed");
    return GE_OK;
}

int ge_ply_vertices(ge_ctx *c, float dw, float dh, ge_ply_vertex *out, uint32_t cap, uint32_t *count) {
    if (!c || !count || (cap && !out)) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    if (!(dw > 0.0f && dh > 0.0f)) return fail(c->err, GE_E_INVALID, "domain_w, domain_h must be > 0");
    *count = 0;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    const State &s = c->b.st[c->b.cur];
    return ply_collect(c->stream, c->n, s.flags, s.x, s.y, nullptr, s.color, s.size, dw, dh, nullptr, cap, out, count,
                       c->err);
}

int ge3_ply_vertices(ge3_ctx *c, ge_ply_vertex *out, uint32_t cap, uint32_t *count) {
    if (!c || !count || (cap && !out)) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    *count = 0;
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    const State3 &s = c->b.st[c->b.cur];
    return ply_collect(c->stream, c->n, s.flags, s.x, s.y, s.z, s.color, s.size, 1.0f, 1.0f, nullptr, cap, out, count,
                       c->err);
}

int ge_ply_format(const ge_ply_vertex *v, uint32_t count, char *text, uint64_t cap, uint64_t *len) {
    if (!len || (count && !v) || (cap && !text)) return GE_E_INVALID;
    const std::string s = ply_text(v, count);
    *len = s.size();
    if (cap) std::memcpy(text, s.data(), (size_t)std::min<uint64_t>(cap, s.size()));
    return GE_OK;
}

int ge_ply_write(ge_ctx *c, float dw, float dh, const char *path) {
    if (!c || !path) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    if (!(dw > 0.0f && dh > 0.0f)) return fail(c->err, GE_E_INVALID, "domain_w, domain_h must be > 0");
    std::vector<ge_ply_vertex> v;
    uint32_t count = 0;
    if (c->n) {
        (void)hipSetDevice(c->device);
        const State &s = c->b.st[c->b.cur];
        if (int r = ply_collect(c->stream, c->n, s.flags, s.x, s.y, nullptr, s.color, s.size, dw, dh, &v, 0, nullptr,
                                &count, c->err))
            return r;
    }
    return write_file(path, ply_text(v.data(), count), c->err);
}

int ge3_ply_write(ge3_ctx *c, const char *path) {
    if (!c || !path) return GE_E_INVALID;
    if (!c->render) return fail(c->err, GE_E_INVALID, "context created without enable_render");
    std::vector<ge_ply_vertex> v;
    uint32_t count = 0;
    if (c->n) {
        (void)hipSetDevice(c->device);
        const State3 &s = c->b.st[c->b.cur];
        if (int r = ply_collect(c->stream, c->n, s.flags, s.x, s.y, s.z, s.color, s.size, 1.0f, 1.0f, &v, 0, nullptr,
                                &count, c->err))
            return r;
    }
    return write_file(path, ply_text(v.data(), count), c->err);
}

void ge_default_galaxy_params(ge_galaxy_params *g) {
    if (!g) return;
    std::memset(g, 0, sizeof *g);
    g->visible = 40000;            // 40000 * particleAmountMultiplier (particlesSpawning.cpp:141)
    g->dark = 12000;               // 12000 * DMAmountMultiplier (particlesSpawning.cpp:200)
    g->outer_radius = 200.0f;      // particlesSpawning.h:28-32
    g->scale_length = 90.0f;
    g->outer_radius_dm = 2000.0f;
    g->radius_core_dm = 3.5f;
    g->mass_visible = 8500000000.0f;   // particlesSpawning.cpp:165
    g->mass_dark = 141600000000.0f;    // particlesSpawning.cpp:222
    g->mass_scatter = 0.75f;           // parameters.h:297
}

}  // extern "C"
