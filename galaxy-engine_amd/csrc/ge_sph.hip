// ge_sph.hip -- SPH neighbour grid and the PCISPH passes of SPH::pcisphSolver
// (GE/include/Physics/SPH.h:623-632, GE/src/Physics/SPH.cpp:98-281) with the
// candidate semantics of NeighborSearchV2AVX2 (GE/include/Particles/
// neighborSearch.h:636-777): a particle's candidates are all particles whose
// grid cell of pos is within one cell (3 x 3) of the query point's cell.
//
// Two search semantics (ge_step_params.sph_search): 0 is the AVX2 grid's
// above; 1 is NeighborSearchV2, the reference's path without AVX2
// (neighborSearch.h:439-599): cell = (int)(p * (1/cell)), ceil(radius/cell)
// cells each way and a filter |p_j - q|^2 <= radius^2 on the grid position.
// Both relations stay symmetric, so the gather forms below hold for both.
//
// Uniform grid instead of the reference's 16384-slot hash: particles are
// radix-sorted by cell over the occupied cell range and first[c] (prefix
// cell table) turns the three cells of one row into one contiguous range.
// Particle fields are packed in cell order, so a neighbour range is read
// contiguously.  The candidate SET is the hash's: the hash only groups cells,
// its queries compare exact cell coordinates.
//
// The reference scatters every pair force to both particles with float
// atomics.  Here each particle gathers both halves itself, so results are
// deterministic:
//   viscosity/cohesion (queries at pos, a symmetric relation): own visits
//     (+F(i,j)) and the neighbours' visits of i (-F(j,i)) from one range;
//   pressure (queries at predPos, not symmetric): own visits from grid A
//     around cell(predPos_i), neighbours' visits from grid B (cells of
//     predPos of visiting particles) around cell(pos_i).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr float kPiF = 3.14159265358979323846f;  // raylib PI, a float
constexpr uint32_t kSentinelCell = (1u << kCellBits) - 1u;
constexpr uint32_t kGridMaxBlocks = 1024;

// candidate-set semantics of one grid
struct Search {
    int mode;        // 0 NeighborSearchV2AVX2, 1 NeighborSearchV2
    float cell, inv; // cell size and the float 1/cell of NeighborSearchV2
    float r2;        // searchRadius^2 (mode 1)
    int R;           // cells searched on each side of the query cell
};

struct SphConst {
    float h, h2, cell, unit, visc, coh, delta, stiff_mult, coeff;
    float vol, spiky, lap;  // kernel constants (SPH.h:38-56)
    float h2hi;             // a bound with sqrt(r2) < h  =>  r2 < h2hi (cheap pre-test)
    float inv_h, inv_vol, coh_norm;  // 1/h, 1/vol, 30/(pi h^2)
    float half_dt;          // fused integrateEnd (0 = none)
    Search S;
};

struct GridView {
    const uint32_t *first;
    const uint32_t *ord;
    const int32_t *frame;  // cx0, cy0, W, H, valid
    const float4 *pk0, *pk1;
    const uint32_t *pkf;
    const float2 *xy;      // position of the candidate, NaN when no pass visits it
};

// a NaN position fails every distance test, so a candidate that takes no part
// costs one 8-byte load
__device__ __forceinline__ float2 lean_pos(bool visits, float x, float y) {
    return visits ? make_float2(x, y) : make_float2(__builtin_nanf(""), __builtin_nanf(""));
}

__device__ __forceinline__ bool is_visitor(uint32_t f) { return (f & GE_FLAG_SPH) && !(f & GE_FLAG_DRAWN); }

// posToCellCoord: NeighborSearchV2AVX2 floor(p / cell) (neighborSearch.h:
// 631-635) or NeighborSearchV2 (int)(p * invCellSize) (neighborSearch.h:
// 472-474), with the float clamped so far-away points cannot overflow the int
// conversion
__device__ __forceinline__ int cell_coord(float p, const Search &S) {
    float c = S.mode ? truncf(p * S.inv) : floorf(p / S.cell);
    c = fminf(fmaxf(c, -1073741824.0f), 1073741824.0f);
    return (int)c;
}

// mode 1's distance filter; dx, dy = grid position - query point
__device__ __forceinline__ bool within(const Search &S, float dx, float dy) {
#pragma clang fp contract(off)
    return !S.mode || dx * dx + dy * dy <= S.r2;
}

// particles of cells (qx-R..qx+R, row) as one range of cell-order positions
__device__ __forceinline__ bool row_range(const int32_t *fr, const uint32_t *first, int qx, int row, int R,
                                          uint32_t &lo, uint32_t &hi) {
    if (!fr[4]) return false;
    const int64_t r = (int64_t)row - fr[1];
    if (r < 0 || r >= fr[3]) return false;
    int64_t c0 = (int64_t)qx - R - fr[0], c1 = (int64_t)qx + R - fr[0];
    if (c1 < 0 || c0 >= fr[2]) return false;
    if (c0 < 0) c0 = 0;
    if (c1 >= fr[2]) c1 = fr[2] - 1;
    const uint64_t base = (uint64_t)r * (uint64_t)fr[2];
    lo = first[base + (uint64_t)c0];
    hi = first[base + (uint64_t)c1 + 1];
    return lo < hi;
}

// particles of cells (c0..c1, row) as one range of cell-order positions
__device__ __forceinline__ bool row_span(const int32_t *fr, const uint32_t *first, int c0, int c1, int row,
                                         uint32_t &lo, uint32_t &hi) {
    if (!fr[4]) return false;
    const int64_t r = (int64_t)row - fr[1];
    if (r < 0 || r >= fr[3]) return false;
    int64_t a = (int64_t)c0 - fr[0], b = (int64_t)c1 - fr[0];
    if (b < 0 || a >= fr[2]) return false;
    if (a < 0) a = 0;
    if (b >= fr[2]) b = fr[2] - 1;
    const uint64_t base = (uint64_t)r * (uint64_t)fr[2];
    lo = first[base + (uint64_t)a];
    hi = first[base + (uint64_t)b + 1];
    return lo < hi;
}

// ---- LDS tiles of candidate rows.  A workgroup's particles are consecutive
// in cell order, so when the query cells of its visitors share one row y,
// the candidates of all of them lie in three contiguous ranges of cell-order
// positions: cells c_min - 1 .. c_max + 1 of rows y - 1, y, y + 1.  Those
// ranges are staged in LDS once per workgroup and every lane reads its
// candidates there.  A workgroup whose queries span rows, or whose rows
// exceed the tile, reads global memory instead; both paths visit the same
// candidates in the same order.  (NeighborSearchV2's wider search, mode 1,
// always reads global memory.)
constexpr uint32_t kTileCand = 384;

struct TileSpan {  // one per workgroup, in LDS
    int cmin, cmax, rmin, rmax;  // query cells of the workgroup's visitors
    uint32_t lo[3];              // first grid position of each staged row
    uint32_t base[4];            // first tile slot of each staged row; base[3] = staged count
    uint32_t ok;
};

__device__ __forceinline__ void span_init(TileSpan &T) {
    T.cmin = 0x7FFFFFFF;
    T.cmax = -0x7FFFFFFF - 1;
    T.rmin = 0x7FFFFFFF;
    T.rmax = -0x7FFFFFFF - 1;
}

__device__ __forceinline__ void span_add(TileSpan &T, int cx, int cy) {
    atomicMin(&T.cmin, cx);
    atomicMax(&T.cmax, cx);
    atomicMin(&T.rmin, cy);
    atomicMax(&T.rmax, cy);
}

// one thread: the staged rows of the span (no visitor: nothing staged)
__device__ __forceinline__ void span_plan(TileSpan &T, const int32_t *fr, const uint32_t *first) {
    const bool one_row = T.rmin == T.rmax;
    uint32_t tot = 0;
    for (int d = 0; d < 3; ++d) {
        uint32_t lo = 0, hi = 0;
        if (!one_row || !row_span(fr, first, T.cmin - 1, T.cmax + 1, T.rmin + d - 1, lo, hi)) lo = hi = 0;
        T.lo[d] = lo;
        T.base[d] = tot;
        tot += hi - lo;
    }
    T.base[3] = tot;
    T.ok = one_row && tot <= kTileCand;
}

// grid position of tile slot e (e < base[3])
__device__ __forceinline__ uint32_t span_pos(const TileSpan &T, uint32_t e) {
    const uint32_t d = e >= T.base[2] ? 2u : (e >= T.base[1] ? 1u : 0u);
    return T.lo[d] + (e - T.base[d]);
}

// The pair passes divide with the hardware reciprocal (v_rcp_f32, within
// 1 ulp) where the reference divides: the pair terms then differ from the
// reference's by a few ulp, far inside the passes' tolerance, at a fraction
// of the issue cost of a correctly rounded division.  Tests that decide
// which pairs interact keep the reference's exact arithmetic.
__device__ __forceinline__ float rcp(float v) { return __builtin_amdgcn_rcpf(v); }

// SPH::smoothingKernelCohesion (SPH.h:58-64)
__device__ __forceinline__ float coh_kernel(float r, const SphConst &K) {
#pragma clang fp contract(off)
    if (r >= K.h) return 0.0f;
    const float q = r * K.inv_h;
    return (1.0f - q) * (0.5f - q) * (0.5f - q) * K.coh_norm;
}

// SPH::spikyKernelDerivative (SPH.h:45-50): float scale times pow(h - r, 2.0f),
// the float overload of pow in the reference's build; (h - r)^2 is formed as
// one correctly rounded float product
__device__ __forceinline__ float spiky_grad(float r, float h, float scale) {
#pragma clang fp contract(off)
    if (r >= h) return 0.0f;
    const float d = h - r;
    return scale * (d * d);
}

// min/max of (x, y) over the included particles, per workgroup
__global__ __launch_bounds__(kBlock) void k_minmax(const float *__restrict__ x, const float *__restrict__ y,
                                                   const uint32_t *__restrict__ flags, uint32_t n, int visitors_only,
                                                   float4 *__restrict__ partials) {
    __shared__ float red[4][kBlock / 64];
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        if (visitors_only && !is_visitor(flags[i])) continue;
        const float px = x[i], py = y[i];
        mnx = glm_min(mnx, px); mny = glm_min(mny, py);
        mxx = glm_max(mxx, px); mxy = glm_max(mxy, py);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mnx = glm_min(mnx, __shfl_xor(mnx, o, 64)); mny = glm_min(mny, __shfl_xor(mny, o, 64));
        mxx = glm_max(mxx, __shfl_xor(mxx, o, 64)); mxy = glm_max(mxy, __shfl_xor(mxy, o, 64));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[0][w] = mnx; red[1][w] = mny; red[2][w] = mxx; red[3][w] = mxy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) {
            mnx = glm_min(mnx, red[0][k]); mny = glm_min(mny, red[1][k]);
            mxx = glm_max(mxx, red[2][k]); mxy = glm_max(mxy, red[3][k]);
        }
        partials[blockIdx.x] = make_float4(mnx, mny, mxx, mxy);
    }
}

// Grid frame from the extremes: the occupied cell range.  One wave: the
// lanes stride the per-workgroup partials, then a wave reduction.
__global__ __launch_bounds__(64) void k_frame(const float4 *__restrict__ partials, uint32_t nparts,
                                              const float *__restrict__ ext_in, Search S, uint32_t cells_cap,
                                              int32_t *__restrict__ frame, uint32_t *__restrict__ counters) {
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    if (ext_in) {
        mnx = ext_in[0]; mny = ext_in[1]; mxx = ext_in[2]; mxy = ext_in[3];
    } else {
        for (uint32_t k = threadIdx.x; k < nparts; k += 64) {
            const float4 v = partials[k];
            mnx = glm_min(mnx, v.x); mny = glm_min(mny, v.y);
            mxx = glm_max(mxx, v.z); mxy = glm_max(mxy, v.w);
        }
        for (int o = 32; o > 0; o >>= 1) {
            mnx = glm_min(mnx, __shfl_xor(mnx, o, 64)); mny = glm_min(mny, __shfl_xor(mny, o, 64));
            mxx = glm_max(mxx, __shfl_xor(mxx, o, 64)); mxy = glm_max(mxy, __shfl_xor(mxy, o, 64));
        }
    }
    if (threadIdx.x != 0) return;
    if (!(mnx <= mxx) || !(mny <= mxy)) {  // no particle in the grid
        frame[0] = 0; frame[1] = 0; frame[2] = 0; frame[3] = 0; frame[4] = 1;
        return;
    }
    const float lim = 536870912.0f;  // keep cell coordinates well inside int
    // the cell function is monotone, so the extremes bound every cell
    const float fx0 = S.mode ? truncf(mnx * S.inv) : floorf(mnx / S.cell);
    const float fy0 = S.mode ? truncf(mny * S.inv) : floorf(mny / S.cell);
    const float fx1 = S.mode ? truncf(mxx * S.inv) : floorf(mxx / S.cell);
    const float fy1 = S.mode ? truncf(mxy * S.inv) : floorf(mxy / S.cell);
    bool ok = fabsf(fx0) < lim && fabsf(fy0) < lim && fabsf(fx1) < lim && fabsf(fy1) < lim;
    const int64_t W = ok ? (int64_t)fx1 - (int64_t)fx0 + 1 : 0;
    const int64_t H = ok ? (int64_t)fy1 - (int64_t)fy0 + 1 : 0;
    ok = ok && W > 0 && H > 0 && W * H <= (int64_t)cells_cap;
    if (!ok) {
        atomicOr(&counters[kCntError], kErrGridRange);
        frame[0] = 0; frame[1] = 0; frame[2] = 0; frame[3] = 0; frame[4] = 0;
        return;
    }
    frame[0] = (int32_t)fx0; frame[1] = (int32_t)fy0; frame[2] = (int32_t)W; frame[3] = (int32_t)H; frame[4] = 1;
}

__global__ __launch_bounds__(kBlock) void k_cell_keys(const float *__restrict__ x, const float *__restrict__ y,
                                                      const uint32_t *__restrict__ flags, uint32_t n, int visitors_only,
                                                      Search S, const int32_t *__restrict__ frame,
                                                      uint64_t *__restrict__ packed) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t key = kSentinelCell;
    if (frame[4] && (!visitors_only || is_visitor(flags[i]))) {
        const int cx = cell_coord(x[i], S), cy = cell_coord(y[i], S);
        key = (uint64_t)(cy - frame[1]) * (uint64_t)frame[2] + (uint64_t)(cx - frame[0]);
    }
    packed[i] = (key << kIdxBits) | i;
}

// first[c] for every cell c in [0, W*H]: the first cell-order position whose
// key is >= c.  Position i fills the cells between its predecessor's key and
// its own, so every entry has exactly one writer.
__global__ __launch_bounds__(kBlock) void k_cell_table(const uint64_t *__restrict__ sorted, uint32_t n,
                                                       const int32_t *__restrict__ frame, uint32_t *__restrict__ first,
                                                       uint32_t *__restrict__ ord) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i > n) return;
    const int64_t ncells = frame[4] ? (int64_t)frame[2] * frame[3] : 0;
    if (i < n) ord[i] = (uint32_t)(sorted[i] & kIdxMask);
    const int64_t ki = i < n ? (int64_t)(sorted[i] >> kIdxBits) : ncells;
    const int64_t kp = i == 0 ? -1 : (int64_t)(sorted[i - 1] >> kIdxBits);
    const int64_t lo = (kp < ncells ? kp : ncells) + 1;
    const int64_t hi = ki < ncells ? ki : ncells;
    for (int64_t c = lo; c <= hi; ++c) first[c] = i;
}

// cell-order packs of grid A: (x, y, vx, vy), (sphMass, visc, cohesion, dens), flags
__global__ __launch_bounds__(kBlock) void k_pack_a(State s, uint32_t n, const uint32_t *__restrict__ ord,
                                                   float4 *__restrict__ pk0, float4 *__restrict__ pk1,
                                                   uint32_t *__restrict__ pkf, float2 *__restrict__ xy) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = ord[k];
    const float x = s.x[i], y = s.y[i];
    const uint32_t f = s.flags[i];
    pk0[k] = make_float4(x, y, s.vx[i], s.vy[i]);
    pk1[k] = make_float4(s.sph_mass[i], s.visc[i], s.cohesion[i], s.dens[i]);
    pkf[k] = f;
    xy[k] = lean_pos(is_visitor(f), x, y);
}

// The pair passes give each particle kSphLanes lanes: lane q of the group
// takes the candidates lo + q, lo + q + kSphLanes, ... of every row, and the
// group adds its partial sums in a fixed butterfly order, so the result is
// deterministic.  At the particle counts of an SPH scene this quadruples the
// waves in flight, which hides the latency of the candidate reads.
constexpr uint32_t kSphLanes = 8;

__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (uint32_t o = 1; o < kSphLanes; o <<= 1) v += __shfl_xor(v, (int)o, 64);
    return v;
}

// SPH::computeViscCohesionForces (SPH.cpp:98-150) in gather form, then the
// predicted position (SPH.cpp:173-181).  pp[k] = (predX, predY, restDens,
// sphMass).  Candidates come from an LDS tile when the workgroup's query
// cells share a row.
__global__ __launch_bounds__(kBlock) void k_visc(State s, uint32_t n, GridView A, SphConst K,
                                                 float2 *__restrict__ sphf, float4 *__restrict__ pp,
                                                 float2 *__restrict__ ppxy) {
#pragma clang fp contract(off)
    __shared__ TileSpan TA;
    __shared__ float2 t_xy[kTileCand], t_v[kTileCand];
    __shared__ float4 t_q[kTileCand];
    __shared__ uint32_t t_f[kTileCand];
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t k = t / kSphLanes, q0 = t % kSphLanes;
    const bool have = k < n;
    const uint32_t i = have ? A.ord[k] : 0u;
    const uint32_t fi = have ? A.pkf[k] : 0u;
    const bool vis = have && is_visitor(fi);
    const float4 own = have ? A.pk0[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 oq = have ? A.pk1[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int cx = vis ? cell_coord(own.x, K.S) : 0, cy = vis ? cell_coord(own.y, K.S) : 0;
    const bool tiles = K.S.mode == 0;  // uniform over the launch
    if (tiles) {
        if (threadIdx.x == 0) span_init(TA);
        __syncthreads();
        if (vis && q0 == 0) span_add(TA, cx, cy);
        __syncthreads();
        if (threadIdx.x == 0) span_plan(TA, A.frame, A.first);
        __syncthreads();
        if (TA.ok)
            for (uint32_t e = threadIdx.x; e < TA.base[3]; e += kBlock) {
                const uint32_t kk = span_pos(TA, e);
                const float4 p0 = A.pk0[kk];
                t_xy[e] = A.xy[kk];
                t_v[e] = make_float2(p0.z, p0.w);
                t_q[e] = A.pk1[kk];
                t_f[e] = A.pkf[kk];
            }
        __syncthreads();
    }
    if (!have) return;
    if (!vis) {
        if (q0 == 0) {
            sphf[i] = make_float2(0.0f, 0.0f);
            pp[k] = make_float4(s.ppx[i], s.ppy[i], s.rest_dens[i], oq.x);
            ppxy[k] = lean_pos(false, 0.0f, 0.0f);
        }
        return;
    }
    const bool ta = tiles && TA.ok;
    const bool fwd = !(fi & GE_FLAG_PINNED);
    const float mI = oq.x * K.unit;
    const float di = oq.w < 0.001f ? 0.001f : oq.w;
    const float aI = K.visc * oq.y * mI * rcp(di) * K.lap;  // j's visit of i: viscosity factor
    const float cI = K.coh * oq.z;
    float sx = 0.0f, sy = 0.0f;
    for (int dr = -K.S.R; dr <= K.S.R; ++dr) {
        uint32_t lo, hi;
        if (!row_range(A.frame, A.first, cx, cy + dr, K.S.R, lo, hi)) continue;
        const uint32_t off = ta ? TA.base[dr + 1] - TA.lo[dr + 1] : 0u;  // tile slot = kk + off
        for (uint32_t kk = lo + q0; kk < hi; kk += kSphLanes) {
            // non-visitors carry NaN and fail both tests
            const float2 q = ta ? t_xy[kk + off] : A.xy[kk];
            const float dx = q.x - own.x, dy = q.y - own.y;
            if (!within(K.S, dx, dy)) continue;
            const float rsq = dx * dx + dy * dy;
            if (!(rsq < K.h2) || kk == k) continue;
            float2 vj;
            float4 qj;
            uint32_t fj;
            if (ta) {
                vj = t_v[kk + off];
                qj = t_q[kk + off];
                fj = t_f[kk + off];
            } else {
                const float4 pj = A.pk0[kk];
                vj = make_float2(pj.z, pj.w);
                qj = A.pk1[kk];
                fj = A.pkf[kk];
            }
            const float r = __builtin_amdgcn_sqrtf(rsq < 1e-6f ? 1e-6f : rsq);
            const float inv_r = rcp(r);
            const float wc = coh_kernel(r, K);
            if (fwd) {  // i's visit of j
                const float mJ = qj.x * K.unit;
                const float dj = qj.w < 0.001f ? 0.001f : qj.w;
                const float a = K.visc * qj.y * mJ * rcp(dj) * K.lap;
                const float vfx = a * (vj.x - own.z);
                const float vfy = a * (vj.y - own.w);
                const float cw = cI * mJ * wc;
                const float cfx = cw * (dx * inv_r);
                const float cfy = cw * (dy * inv_r);
                sx = sx + (vfx + cfx);
                sy = sy + (vfy + cfy);
            }
            if (!(fj & GE_FLAG_PINNED)) {  // j's visit of i, scattered back to i
                const float ex = -dx, ey = -dy;  // pos_i - pos_j
                const float vfx = aI * (own.z - vj.x);
                const float vfy = aI * (own.w - vj.y);
                const float cw = K.coh * qj.z * mI * wc;
                const float cfx = cw * (ex * inv_r);
                const float cfy = cw * (ey * inv_r);
                sx = sx - (vfx + cfx);
                sy = sy - (vfy + cfy);
            }
        }
    }
    sx = group_sum(sx);
    sy = group_sum(sy);
    if (q0 != 0) return;
    sphf[i] = make_float2(sx, sy);
    const float px = own.x + (sx / oq.x) * K.coeff;
    const float py = own.y + (sy / oq.x) * K.coeff;
    s.ppx[i] = px;
    s.ppy[i] = py;
    pp[k] = make_float4(px, py, s.rest_dens[i], oq.x);
    ppxy[k] = make_float2(px, py);
}

// predicted density and pressure (SPH.cpp:183-219); pd[k] = (press,
// predDens).  Candidates from an LDS tile as in k_visc (queries at predPos).
__global__ __launch_bounds__(kBlock) void k_density(State s, uint32_t n, GridView A, const float4 *__restrict__ pp,
                                                    const float2 *__restrict__ ppxy, SphConst K,
                                                    float2 *__restrict__ pd) {
#pragma clang fp contract(off)
    __shared__ TileSpan TA;
    __shared__ float2 t_xy[kTileCand], t_rm[kTileCand];
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t k = t / kSphLanes, q0 = t % kSphLanes;
    const bool have = k < n;
    const uint32_t i = have ? A.ord[k] : 0u;
    const uint32_t fi = have ? A.pkf[k] : 0u;
    const bool vis = have && is_visitor(fi);
    const float4 me = vis ? pp[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int cx = vis ? cell_coord(me.x, K.S) : 0, cy = vis ? cell_coord(me.y, K.S) : 0;
    const bool tiles = K.S.mode == 0;  // uniform over the launch
    if (tiles) {
        if (threadIdx.x == 0) span_init(TA);
        __syncthreads();
        if (vis && q0 == 0) span_add(TA, cx, cy);
        __syncthreads();
        if (threadIdx.x == 0) span_plan(TA, A.frame, A.first);
        __syncthreads();
        if (TA.ok)
            for (uint32_t e = threadIdx.x; e < TA.base[3]; e += kBlock) {
                const uint32_t kk = span_pos(TA, e);
                const float4 pj = pp[kk];
                t_xy[e] = ppxy[kk];
                t_rm[e] = make_float2(pj.z, pj.w);
            }
        __syncthreads();
    }
    if (!have) return;
    if (!vis) {
        if (q0 == 0) {
            s.press[i] = 0.0f;  // press is cleared for every particle (SPH.cpp:162-165)
            pd[k] = make_float2(0.0f, s.pred_dens[i]);
        }
        return;
    }
    const bool ta = tiles && TA.ok;
    float dens = 0.0f;
    for (int dr = -K.S.R; dr <= K.S.R; ++dr) {
        uint32_t lo, hi;
        if (!row_range(A.frame, A.first, cx, cy + dr, K.S.R, lo, hi)) continue;
        const uint32_t off = ta ? TA.base[dr + 1] - TA.lo[dr + 1] : 0u;
        for (uint32_t kk = lo + q0; kk < hi; kk += kSphLanes) {
            // predicted positions, NaN for non-visitors (they fail the test)
            const float2 q = ta ? t_xy[kk + off] : ppxy[kk];
            const float dx = me.x - q.x, dy = me.y - q.y;
            const float rrsq = dx * dx + dy * dy;
            if (!(rrsq < K.h * K.h)) continue;
            if (K.S.mode) {  // filter on the candidate's grid position (pos)
                const float2 gj = A.xy[kk];
                if (!within(K.S, gj.x - me.x, gj.y - me.y)) continue;
            }
            float2 rm;  // (restDens, sphMass) of the candidate
            if (ta) {
                rm = t_rm[kk + off];
            } else {
                const float4 pj = pp[kk];
                rm = make_float2(pj.z, pj.w);
            }
            const float rr = __builtin_amdgcn_sqrtf(rrsq);
            const float mJ = rm.y * K.unit;
            const float rho0 = 0.5f * (me.z + rm.x);
            const float w = (K.h - rr) * (K.h - rr) * K.inv_vol;
            dens = dens + mJ * w * rcp(rho0);
        }
    }
    dens = group_sum(dens);
    if (q0 != 0) return;
    float pt = K.delta * (dens - me.z);
    if (pt < 0.0f) pt = 0.0f;
    const float pr = 0.0f + pt * s.stiff[i] * K.stiff_mult;
    s.pred_dens[i] = dens;
    s.press_tmp[i] = pt;
    s.press[i] = pr;
    pd[k] = make_float2(pr, dens);
}

// grid B packs: (predX, predY, press, predDens), (sphMass, -, -, -)
__global__ __launch_bounds__(kBlock) void k_pack_b(State s, uint32_t n, const uint32_t *__restrict__ ord,
                                                   float4 *__restrict__ pk0, float4 *__restrict__ pk1,
                                                   float2 *__restrict__ xy) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = ord[k];
    const float x = s.ppx[i], y = s.ppy[i];
    pk0[k] = make_float4(x, y, s.press[i], s.pred_dens[i]);
    pk1[k] = make_float4(s.sph_mass[i], 0.0f, 0.0f, 0.0f);
    xy[k] = make_float2(x, y);  // grid B holds visitors only
}

// pressure force of one visit (SPH.cpp:233-256); dx, dy = predPos_visitor - predPos_other
__device__ __forceinline__ void pressure_pair(float dx, float dy, float rr, float mi, float mj, float pi_, float pj_,
                                              float di, float dj, const SphConst &K, float &fx, float &fy) {
#pragma clang fp contract(off)
    const float gw = spiky_grad(rr, K.h, K.spiky);
    const float inv_rr = rcp(rr);
    const float nx = dx * inv_rr, ny = dy * inv_rr;
    const float avgP = 0.5f * (pi_ + pj_);
    const float avgD = 0.5f * (di + dj);
    float mag = -(mi * K.unit + mj * K.unit) * avgP * rcp(avgD < 0.01f ? 0.01f : avgD);
    const float smax = mi < mj ? mj : mi;
    const float smin = mj < mi ? mj : mi;
    const float sc8 = 8.0f * smin * rcp(smax);  // 8 / (smax / smin)
    mag = mag * (sc8 < 1.0f ? sc8 : 1.0f);
    fx = mag * gw * nx;
    fy = mag * gw * ny;
}

// pressure forces (SPH.cpp:221-264), pressF and acc (SPH.cpp:271-280), then
// the integrateEnd kick (physics.cpp:1114-1125) when the frame asks for it.
// Grid-A candidates (queries at predPos) and grid-B candidates (queries at
// pos) each come from an LDS tile when the workgroup's queries allow.
__global__ __launch_bounds__(kBlock) void k_pressure(State s, uint32_t n, GridView A, const float4 *__restrict__ pp,
                                                     const float2 *__restrict__ ppxy, const float2 *__restrict__ pd,
                                                     GridView B, SphConst K, const float2 *__restrict__ sphf) {
#pragma clang fp contract(off)
    __shared__ TileSpan TA, TB;
    __shared__ float2 a_xy[kTileCand], a_pd[kTileCand];
    __shared__ float a_m[kTileCand];
    __shared__ float2 b_xy[kTileCand], b_pq[kTileCand];
    __shared__ float b_m[kTileCand];
    __shared__ uint32_t b_ord[kTileCand];
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t k = t / kSphLanes, q0 = t % kSphLanes;
    const bool have = k < n;
    const uint32_t i = have ? A.ord[k] : 0u;
    const uint32_t fi = have ? A.pkf[k] : 0u;
    const bool vis = have && is_visitor(fi);
    float4 me = make_float4(0.f, 0.f, 0.f, 0.f), mypos = me;
    float2 mp = make_float2(0.f, 0.f);
    int qx = 0, qy = 0, px = 0, py = 0;
    if (vis) {
        me = pp[k];
        mp = pd[k];
        mypos = A.pk0[k];
        qx = cell_coord(me.x, K.S);
        qy = cell_coord(me.y, K.S);
        px = cell_coord(mypos.x, K.S);
        py = cell_coord(mypos.y, K.S);
    }
    const bool tiles = K.S.mode == 0;  // uniform over the launch
    if (tiles) {
        if (threadIdx.x == 0) {
            span_init(TA);
            span_init(TB);
        }
        __syncthreads();
        if (vis && q0 == 0) {
            span_add(TA, qx, qy);
            span_add(TB, px, py);
        }
        __syncthreads();
        if (threadIdx.x == 0) span_plan(TA, A.frame, A.first);
        if (threadIdx.x == 64) span_plan(TB, B.frame, B.first);
        __syncthreads();
        if (TA.ok)
            for (uint32_t e = threadIdx.x; e < TA.base[3]; e += kBlock) {
                const uint32_t kk = span_pos(TA, e);
                a_xy[e] = ppxy[kk];
                a_m[e] = pp[kk].w;
                a_pd[e] = pd[kk];
            }
        if (TB.ok)
            for (uint32_t e = threadIdx.x; e < TB.base[3]; e += kBlock) {
                const uint32_t kb = span_pos(TB, e);
                const float4 q = B.pk0[kb];
                b_xy[e] = make_float2(q.x, q.y);
                b_pq[e] = make_float2(q.z, q.w);
                b_m[e] = B.pk1[kb].x;
                b_ord[e] = B.ord[kb];
            }
        __syncthreads();
    }
    if (!have) return;
    float2 f = sphf[i];
    float gx = 0.0f, gy = 0.0f;  // this lane's share of the pair forces
    if (vis) {
        const bool ta = tiles && TA.ok, tb = tiles && TB.ok;
        // i's visits: candidates by pos cell around cell(predPos_i)
        for (int dr = -K.S.R; dr <= K.S.R; ++dr) {
            uint32_t lo, hi;
            if (!row_range(A.frame, A.first, qx, qy + dr, K.S.R, lo, hi)) continue;
            const uint32_t off = ta ? TA.base[dr + 1] - TA.lo[dr + 1] : 0u;  // tile slot = kk + off
            for (uint32_t kk = lo + q0; kk < hi; kk += kSphLanes) {
                const float2 qp = ta ? a_xy[kk + off] : ppxy[kk];  // NaN for non-visitors: fails the test
                const float dx = me.x - qp.x, dy = me.y - qp.y;
                const float rsq = dx * dx + dy * dy;
                if (!(rsq < K.h2hi)) continue;  // certainly outside h; the exact test follows
                const float rr = sqrtf(rsq);
                if (!(rr >= 1e-5f && rr < K.h) || kk == k) continue;
                if (K.S.mode) {
                    const float2 gj = A.xy[kk];
                    if (!within(K.S, gj.x - me.x, gj.y - me.y)) continue;
                }
                const float mj = ta ? a_m[kk + off] : pp[kk].w;
                const float2 q = ta ? a_pd[kk + off] : pd[kk];
                float fx, fy;
                pressure_pair(dx, dy, rr, me.w, mj, mp.x, q.x, mp.y, q.y, K, fx, fy);
                gx = gx + fx;
                gy = gy + fy;
            }
        }
        // visits of i by others: predPos cells around cell(pos_i)
        for (int dr = -K.S.R; dr <= K.S.R; ++dr) {
            uint32_t lo, hi;
            if (!row_range(B.frame, B.first, px, py + dr, K.S.R, lo, hi)) continue;
            const uint32_t off = tb ? TB.base[dr + 1] - TB.lo[dr + 1] : 0u;
            for (uint32_t kb = lo + q0; kb < hi; kb += kSphLanes) {
                const float2 qp = tb ? b_xy[kb + off] : B.xy[kb];
                const float dx = qp.x - me.x, dy = qp.y - me.y;
                const float rsq = dx * dx + dy * dy;
                if (!(rsq < K.h2hi)) continue;
                const float rr = sqrtf(rsq);
                if (!(rr >= 1e-5f && rr < K.h)) continue;
                // j's query at predPos_j meets i at pos_i
                if (!within(K.S, mypos.x - qp.x, mypos.y - qp.y)) continue;
                if ((tb ? b_ord[kb + off] : B.ord[kb]) == i) continue;
                float2 pq;
                float mj;
                if (tb) {
                    pq = b_pq[kb + off];
                    mj = b_m[kb + off];
                } else {
                    const float4 pj = B.pk0[kb];
                    pq = make_float2(pj.z, pj.w);
                    mj = B.pk1[kb].x;
                }
                float fx, fy;
                pressure_pair(dx, dy, rr, mj, me.w, pq.x, mp.x, pq.y, mp.y, K, fx, fy);
                gx = gx - fx;
                gy = gy - fy;
            }
        }
    }
    gx = group_sum(gx);
    gy = group_sum(gy);
    if (q0 != 0) return;
    f.x = f.x + gx;
    f.y = f.y + gy;
    const float sm = s.sph_mass[i];
    const float pfx = f.x / sm, pfy = f.y / sm;
    s.pfx[i] = pfx;
    s.pfy[i] = pfy;
    if (!(fi & GE_FLAG_PINNED)) {
        const float ax = s.ax[i] + pfx, ay = s.ay[i] + pfy;
        s.ax[i] = ax;
        s.ay[i] = ay;
        if (K.half_dt != 0.0f) {
            s.vx[i] = s.vx[i] + ax * K.half_dt;
            s.vy[i] = s.vy[i] + ay * K.half_dt;
        }
    }
}

// ---- candidate lists for the parity hook; counts (cnt) and lists (out),
// optionally skipping particles whose flags match skip_mask (count 0, or
// left as it was with keep_skipped)
__global__ __launch_bounds__(kBlock) void k_cand_count(const float *__restrict__ qx, const float *__restrict__ qy,
                                                       uint32_t n, Search S, const int32_t *__restrict__ frame,
                                                       const uint32_t *__restrict__ first,
                                                       const float4 *__restrict__ pk0,
                                                       const uint32_t *__restrict__ flags, uint32_t skip_mask,
                                                       int keep_skipped, uint32_t *__restrict__ cnt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (flags && (flags[i] & skip_mask)) {
        if (!keep_skipped) cnt[i] = 0;
        return;
    }
    const float x = qx[i], y = qy[i];
    const int cx = cell_coord(x, S), cy = cell_coord(y, S);
    uint32_t c = 0;
    for (int dr = -S.R; dr <= S.R; ++dr) {
        uint32_t lo, hi;
        if (!row_range(frame, first, cx, cy + dr, S.R, lo, hi)) continue;
        if (!S.mode) {
            c += hi - lo;
            continue;
        }
        for (uint32_t kk = lo; kk < hi; ++kk) {
            const float4 g = pk0[kk];
            c += within(S, g.x - x, g.y - y) ? 1u : 0u;
        }
    }
    cnt[i] = c;
}

__global__ __launch_bounds__(kBlock) void k_cand_fill(const float *__restrict__ qx, const float *__restrict__ qy,
                                                      uint32_t n, Search S, const int32_t *__restrict__ frame,
                                                      const uint32_t *__restrict__ first,
                                                      const float4 *__restrict__ pk0, const uint32_t *__restrict__ ord,
                                                      const int64_t *__restrict__ off, uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float x = qx[i], y = qy[i];
    const int cx = cell_coord(x, S), cy = cell_coord(y, S);
    int64_t w = off[i];
    for (int dr = -S.R; dr <= S.R; ++dr) {
        uint32_t lo, hi;
        if (!row_range(frame, first, cx, cy + dr, S.R, lo, hi)) continue;
        for (uint32_t kk = lo; kk < hi; ++kk) {
            const float4 g = pk0[kk];
            if (within(S, g.x - x, g.y - y)) out[w++] = ord[kk];
        }
    }
}

// ---- Physics::createConstraints' automatic branch (physics.cpp:654-766):
// particles with hasSolidified link to neighbour candidates of their position
// ----------------------------------------------------------------------------

constexpr int kCcBlock = 1024;

__device__ __forceinline__ const ge_material *cc_material(const ge_material *mats, uint32_t nmat, uint32_t id) {
    for (uint32_t k = 0; k < nmat; ++k)
        if (mats[k].id == id) return &mats[k];
    return nullptr;
}

// Calls f(j, dist_sq) for every candidate j of slot i that the automatic
// branch admits, in the reference's visiting order: cells dx outer, dy inner
// (neighborSearch.h:730-731, 559-560), ascending slot inside a cell (the grid
// orders a cell by slot; NeighborSearchV2's counting sort does the same).
template <class F>
__device__ __forceinline__ void cc_visit(uint32_t i, const State &s, const GridView &A, const Search &S,
                                         const ge_material *mats, uint32_t nmat, F &&f) {
#pragma clang fp contract(off)
    const float xi = s.x[i], yi = s.y[i];
    const uint32_t idi = s.id[i];
    const ge_material *mi = s.label ? cc_material(mats, nmat, s.label[i]) : nullptr;
    const int qx = cell_coord(xi, S), qy = cell_coord(yi, S);
    const int32_t *fr = A.frame;
    if (!fr[4]) return;
    for (int dx = -S.R; dx <= S.R; ++dx) {
        const int64_t col = (int64_t)qx + dx - fr[0];
        if (col < 0 || col >= fr[2]) continue;
        for (int dy = -S.R; dy <= S.R; ++dy) {
            const int64_t row = (int64_t)qy + dy - fr[1];
            if (row < 0 || row >= fr[3]) continue;
            const uint64_t cidx = (uint64_t)row * (uint64_t)fr[2] + (uint64_t)col;
            const uint32_t lo = A.first[cidx], hi = A.first[cidx + 1];
            for (uint32_t kk = lo; kk < hi; ++kk) {
                const float4 g = A.pk0[kk];
                if (!within(S, g.x - xi, g.y - yi)) continue;  // NeighborSearchV2's radius filter
                const uint32_t j = A.ord[kk];
                if (j == i) continue;
                const float ex = g.x - xi, ey = g.y - yi;
                const float dist_sq = ex * ex + ey * ey;  // glm::dot (physics.cpp:706)
                if (dist_sq > 12.0f) continue;
                const ge_material *mj = s.label ? cc_material(mats, nmat, s.label[j]) : nullptr;
                if (mi && mj && mi->cold_point == 0.0f && mj->cold_point != 0.0f) continue;
                if (!(idi < s.id[j])) continue;
                f(j, dist_sq, mi, mj);
            }
        }
    }
}

// per slot: the number of constraints it creates; hasSolidified is consumed
__global__ __launch_bounds__(kCcBlock) void k_cc_count(State s, uint32_t n, GridView A, Search S,
                                                       const ge_material *__restrict__ mats, uint32_t nmat,
                                                       uint32_t *__restrict__ cnt, uint32_t *__restrict__ bsum) {
    __shared__ uint32_t red[kCcBlock / 64];
    const uint32_t i = blockIdx.x * kCcBlock + threadIdx.x;
    uint32_t k = 0;
    if (i < n && (s.flags[i] & GE_FLAG_SOLIDIFIED)) {
        cc_visit(i, s, A, S, mats, nmat, [&](uint32_t, float, const ge_material *, const ge_material *) { ++k; });
        s.flags[i] &= ~GE_FLAG_SOLIDIFIED;  // pi.hasSolidified = false (physics.cpp:693)
    }
    if (i < n) cnt[i] = k;
    uint32_t w = k;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int q = 0; q < kCcBlock / 64; ++q) t += red[q];
        bsum[blockIdx.x] = t;
    }
}

// one workgroup: exclusive scan of the nb block sums in place; the created
// total and the first free constraint index go to tot[0], tot[1], and the
// count grows by what fits (an overflow raises kErrConsCap)
__global__ __launch_bounds__(1024) void k_cc_scan(uint32_t *__restrict__ bsum, uint32_t nb,
                                                  uint32_t *__restrict__ ccount, uint32_t cap,
                                                  uint32_t *__restrict__ tot, uint32_t *__restrict__ err) {
    __shared__ uint32_t sh[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t t = base + threadIdx.x;
        const uint32_t v = t < nb ? bsum[t] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
            const uint32_t a = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        if (t < nb) bsum[t] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t first = ccount[0];
        const uint64_t want = (uint64_t)first + carry;
        tot[0] = carry;
        tot[1] = first;
        if (want > cap) atomicOr(err, kErrConsCap);
        ccount[0] = want > cap ? cap : (uint32_t)want;
    }
}

// each creating slot writes its constraints at first + its exclusive offset
__global__ __launch_bounds__(kCcBlock) void k_cc_write(State s, uint32_t n, GridView A, Search S,
                                                       const ge_material *__restrict__ mats, uint32_t nmat,
                                                       const uint32_t *__restrict__ cnt,
                                                       const uint32_t *__restrict__ bsum,
                                                       const uint32_t *__restrict__ tot, uint32_t cap,
                                                       ge_constraint *__restrict__ cons) {
#pragma clang fp contract(off)
    __shared__ uint32_t sh[kCcBlock];
    const uint32_t i = blockIdx.x * kCcBlock + threadIdx.x;
    const uint32_t k = i < n ? cnt[i] : 0u;
    sh[threadIdx.x] = k;
    __syncthreads();
    for (uint32_t o = 1; o < kCcBlock; o <<= 1) {
        const uint32_t a = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += a;
        __syncthreads();
    }
    if (k == 0) return;
    uint64_t at = (uint64_t)tot[1] + bsum[blockIdx.x] + sh[threadIdx.x] - k;
    const uint32_t idi = s.id[i];
    cc_visit(i, s, A, S, mats, nmat, [&](uint32_t j, float dist_sq, const ge_material *mi, const ge_material *mj) {
        if (at < cap) {
            ge_constraint r;
            r.id1 = idi;
            r.id2 = s.id[j];
            r.rest_length = r.original_length = sqrtf(dist_sq);  // glm::distance
            const bool both = mi && mj;
            r.stiffness = both ? mi->constraint_stiffness : 60.0f;
            r.resistance = both ? (mi->constraint_resistance + mj->constraint_resistance) * 0.5f : 0.6f;
            r.plasticity_point = both ? (mi->constraint_plastic_point + mj->constraint_plastic_point) * 0.5f : 0.6f;
            r.displacement = 0.0f;
            r.is_broken = 0;
            r.is_plastic = 0;
            r.pad[0] = r.pad[1] = 0;
            cons[at] = r;
        }
        ++at;
    });
}

Search make_search(int mode, float cell, float radius) {
    Search S;
    S.mode = mode;
    S.cell = cell;
    S.inv = 1.0f / cell;
    S.r2 = radius * radius;
    S.R = mode ? (int)std::ceil(radius * S.inv) : 1;  // neighborSearch.h:552
    return S;
}

SphConst make_const(const ge_step_params *p, float half_dt) {
    SphConst K;
    K.h = p->sph_h;
    K.h2 = p->sph_h * p->sph_h;
    K.cell = p->sph_cell;
    K.unit = p->sph_unit_mass;
    K.visc = p->viscosity;
    K.coh = p->cohesion_coef;
    K.delta = p->delta;
    K.stiff_mult = p->stiff_mult;
    K.coeff = 0.5f * p->dt * p->dt;  // predictionCoeff (SPH.cpp:156)
    // SPH.h:38-56 in float: pow(float, float) is std::pow's float overload in
    // the reference's build, and PI is raylib's float constant
    K.vol = (kPiF * std::pow(K.h, 4.0f)) / 6.0f;
    K.spiky = -45.0f / (kPiF * std::pow(K.h, 6.0f));
    K.lap = 45.0f / (kPiF * std::pow(K.h, 6.0f));
    K.half_dt = half_dt;
    // sqrt is correctly rounded and h a float, so sqrt(r2) < h implies r2 < h^2
    // exactly; the float bound sits a little above h^2
    K.h2hi = std::nextafter((float)((double)K.h * K.h * (1.0 + 1e-6)), INFINITY);
    K.inv_h = 1.0f / K.h;
    K.inv_vol = 1.0f / K.vol;
    K.coh_norm = 30.0f / (kPiF * K.h * K.h);
    K.S = make_search(p->sph_search, p->sph_cell, p->sph_search_radius);
    return K;
}

GridView view(const Grid &g) { return GridView{g.first, g.ord, g.frame, g.pk0, g.pk1, g.pkf, g.xy}; }

uint32_t grid_blocks(uint32_t n) {
    uint32_t b = blocks_for(n, kBlock);
    return b > kGridMaxBlocks ? kGridMaxBlocks : b;
}

// sort by cell and build first[] / ord[] for one grid
void build_grid(ge_ctx *c, Grid &g) {
    const uint32_t n = c->n;
    radix_sort_packed(c, g.keys, g.keys_alt, n, kIdxBits, kIdxBits + kCellBits);
    hipLaunchKernelGGL(k_cell_table, dim3(blocks_for((uint64_t)n + 1, kBlock)), dim3(kBlock), 0, c->stream, g.keys, n,
                       g.frame, g.first, g.ord);
}

}  // namespace

// NeighborSearchV2AVX2::newGridAVX2 (neighborSearch.h:636-720): grid A over
// pos, using the extremes stored with the frame's bounding box.
int sph_new_grid(ge_ctx *c, const ge_step_params *p) {
    Grid &g = c->b.grid_pos;
    const State &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const Search S = make_search(p->sph_search, p->sph_cell, p->sph_search_radius);
    hipLaunchKernelGGL(k_frame, dim3(1), dim3(64), 0, c->stream, nullptr, 0u, c->b.bb + 4, S, g.cells_cap, g.frame,
                       c->b.counters);
    hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, s.flags, n, 0, S,
                       g.frame, g.keys);
    build_grid(c, g);
    hipLaunchKernelGGL(k_pack_a, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, s, n, g.ord, g.pk0, g.pk1,
                       g.pkf, g.xy);
    c->have_grid = true;
    c->sph_cell = p->sph_cell;
    c->sph_search = p->sph_search;
    c->sph_radius = p->sph_search_radius;
    return GE_OK;
}

// SPH::PCISPH (SPH.cpp:152-281), one iteration
int sph_pcisph(ge_ctx *c, const ge_step_params *p, float end_kick_dt) {
    Grid &A = c->b.grid_pos;
    Grid &B = c->b.grid_pred;
    const State &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const SphConst K = make_const(p, end_kick_dt * 0.5f);
    const uint32_t nb = blocks_for(n, kBlock);
    const uint32_t nbl = blocks_for((uint64_t)n * kSphLanes, kBlock);  // the pair passes: kSphLanes per particle
    float2 *sphf = reinterpret_cast<float2 *>(c->b.sphf);
    // pp (predX, predY, restDens, sphMass) and pd (press, predDens) in grid-A order
    float4 *ppA = reinterpret_cast<float4 *>(c->b.sph_pp);
    float2 *pdA = reinterpret_cast<float2 *>(c->b.sph_pd);
    float2 *ppxy = reinterpret_cast<float2 *>(c->b.sph_ppxy);
    hipLaunchKernelGGL(k_visc, dim3(nbl), dim3(kBlock), 0, c->stream, s, n, view(A), K, sphf, ppA, ppxy);
    // grid B over the predicted positions of visiting particles
    const uint32_t mb = grid_blocks(n);
    hipLaunchKernelGGL(k_minmax, dim3(mb), dim3(kBlock), 0, c->stream, s.ppx, s.ppy, s.flags, n, 1,
                       reinterpret_cast<float4 *>(c->b.partials2));
    hipLaunchKernelGGL(k_frame, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<const float4 *>(c->b.partials2), mb,
                       nullptr, K.S, B.cells_cap, B.frame, c->b.counters);
    hipLaunchKernelGGL(k_cell_keys, dim3(nb), dim3(kBlock), 0, c->stream, s.ppx, s.ppy, s.flags, n, 1, K.S, B.frame,
                       B.keys);
    build_grid(c, B);
    hipLaunchKernelGGL(k_density, dim3(nbl), dim3(kBlock), 0, c->stream, s, n, view(A), ppA, ppxy, K, pdA);
    hipLaunchKernelGGL(k_pack_b, dim3(nb), dim3(kBlock), 0, c->stream, s, n, B.ord, B.pk0, B.pk1, B.xy);
    hipLaunchKernelGGL(k_pressure, dim3(nbl), dim3(kBlock), 0, c->stream, s, n, view(A), ppA, ppxy, pdA, view(B), K,
                       sphf);
    return GE_OK;
}

// queryNeighborsAVX2 candidate sets of grid A at pos (which 0) or predPos (1)
int sph_neighbors(ge_ctx *c, int which, int64_t *off, uint32_t *idx, int64_t cap, int64_t *total) {
    const Grid &g = c->b.grid_pos;
    const State &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const float *qx = which ? s.ppx : s.x;
    const float *qy = which ? s.ppy : s.y;
    uint32_t *dcnt = nullptr;
    int64_t *doff = nullptr;
    uint32_t *dout = nullptr;
    // the semantics of the grid are the ones it was built with
    const Search S = make_search(c->sph_search, c->sph_cell, c->sph_radius);
    GE_HIP_CHECK(c, hipMalloc(&dcnt, sizeof(uint32_t) * (n + 1)));
    hipLaunchKernelGGL(k_cand_count, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, qx, qy, n, S, g.frame,
                       g.first, g.pk0, nullptr, 0u, 0, dcnt);
    std::vector<uint32_t> hc(n);
    GE_HIP_CHECK(c, hipMemcpyAsync(hc.data(), dcnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    off[0] = 0;
    for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + hc[i];
    *total = off[n];
    if (!idx) {
        (void)hipFree(dcnt);
        return GE_OK;
    }
    if (cap < off[n]) {
        (void)hipFree(dcnt);
        c->err = "neighbour buffer too small";
        return GE_E_CAPACITY;
    }
    GE_HIP_CHECK(c, hipMalloc(&doff, sizeof(int64_t) * (n + 1)));
    GE_HIP_CHECK(c, hipMalloc(&dout, sizeof(uint32_t) * (off[n] > 0 ? off[n] : 1)));
    GE_HIP_CHECK(c, hipMemcpyAsync(doff, off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_cand_fill, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, qx, qy, n, S, g.frame,
                       g.first, g.pk0, g.ord, doff, dout);
    GE_HIP_CHECK(c, hipMemcpyAsync(idx, dout, sizeof(uint32_t) * off[n], hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(dcnt);
    (void)hipFree(doff);
    (void)hipFree(dout);
    for (uint32_t i = 0; i < n; ++i) std::sort(idx + off[i], idx + off[i + 1]);
    return GE_OK;
}

// NeighborSearchV2::newGrid + neighborAmount (neighborSearch.h:481-598)
int sph_neighbor_counts(ge_ctx *c, const ge_step_params *p, uint32_t *counts) {
    ge_step_params q = *p;
    q.sph_search = 1;
    if (int r = sph_new_grid(c, &q)) return r;
    const Grid &g = c->b.grid_pos;
    const State &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    const Search S = make_search(1, q.sph_cell, q.sph_search_radius);
    uint32_t *dcnt = reinterpret_cast<uint32_t *>(c->b.sphf);  // SPH force scratch: 2 floats per particle
    hipLaunchKernelGGL(k_cand_count, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, n, S, g.frame,
                       g.first, g.pk0, s.flags, GE_FLAG_PINNED | GE_FLAG_DRAWN | GE_FLAG_DARK, 0, dcnt);
    GE_HIP_CHECK(c, hipMemcpyAsync(counts, dcnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GE_OK;
}

// neighborAmount of NeighborSearchV2AVX2 (search 0, neighborSearch.h:779-796)
// or NeighborSearchV2 (search 1, neighborSearch.h:585-598) into out[] (device),
// leaving the counts of pinned, drawn and dark-matter particles as they were.
// Rebuilds grid A with these search settings; a later PCISPH needs a new grid.
int sph_count_into(ge_ctx *c, int search, float cell, float radius, int32_t *out) {
    ge_step_params q{};
    q.sph_search = search;
    q.sph_cell = cell;
    q.sph_search_radius = radius;
    if (int r = sph_new_grid(c, &q)) return r;
    const Grid &g = c->b.grid_pos;
    const State &s = c->b.st[c->b.cur];
    const Search S = make_search(search, cell, radius);
    hipLaunchKernelGGL(k_cand_count, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s.x, s.y, c->n, S,
                       g.frame, g.first, g.pk0, s.flags, GE_FLAG_PINNED | GE_FLAG_DRAWN | GE_FLAG_DARK, 1,
                       reinterpret_cast<uint32_t *>(out));
    c->have_grid = false;
    return GE_OK;
}

}  // namespace ge

namespace ge {

// Physics::createConstraints' automatic branch on the current state, over
// grid A (built by the caller on the current positions).  *created gets the
// number of constraints appended (a host read of one word).
int sph_create_constraints(ge_ctx *c, const ge_step_params *p, uint32_t *created) {
    *created = 0;
    const uint32_t n = c->n;
    if (n == 0 || c->cons_cap == 0) return GE_OK;
    Buffers &b = c->b;
    const State &s = b.st[b.cur];
    const Search S = make_search(p->sph_search, p->sph_cell, p->sph_search_radius);
    const GridView A = view(b.grid_pos);
    const ge_material *mats = c->thermal ? b.materials : nullptr;
    const uint32_t nmat = c->thermal ? c->nmat : 0u;
    const uint32_t nb = blocks_for(n, kCcBlock);
    uint32_t *cnt = b.ccreate, *bsum = b.ccreate + c->cap, *tot = bsum + blocks_for(c->cap, kCcBlock);
    hipLaunchKernelGGL(k_cc_count, dim3(nb), dim3(kCcBlock), 0, c->stream, s, n, A, S, mats, nmat, cnt, bsum);
    hipLaunchKernelGGL(k_cc_scan, dim3(1), dim3(1024), 0, c->stream, bsum, nb, b.ccount, c->cons_cap, tot,
                       b.counters + kCntError);
    hipLaunchKernelGGL(k_cc_write, dim3(nb), dim3(kCcBlock), 0, c->stream, s, n, A, S, mats, nmat, cnt, bsum, tot,
                       c->cons_cap, b.cons);
    uint32_t h = 0;
    if (hipMemcpyAsync(&h, tot, sizeof h, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "created constraint count copy failed";
        return GE_E_HIP;
    }
    *created = h;
    if (h) {
        const uint64_t bound = (uint64_t)c->cons_bound + h;
        c->cons_bound = bound > c->cons_cap ? c->cons_cap : (uint32_t)bound;
        c->cons_colored = false;  // new constraints: index map and colours again
    }
    return GE_OK;
}

}  // namespace ge
