// ge_multipole.hip -- node multipole moments and opening radii for the
// accuracy modes (ge_step_params.accuracy = 1: quadrupole, 2: up to the
// hexadecapole).  Neither has a counterpart in the reference, whose walk uses
// monopoles only (physics.cpp:250-259).
//
// Every node is a contiguous range of the Morton-sorted particles.  Its raw
// moments S_ab = sum m (x - c)^a (y - c)^b (a + b <= 4) are summed in FP64
// directly about the node's float centre of mass c, so the offsets are at
// most the node's own size and no term cancels against a larger one:
//   - a node of at most kMpThreadMax particles: one thread;
//   - up to kMpWaveMax particles: one wave, lanes striding the range, then a
//     fixed-order wave reduction;
//   - larger: pieces of kMpWaveMax particles, one wave each, summed per node
//     in piece order.
// The lists of wave and piece nodes are filled with atomics, but every
// node's sums have a fixed order, so the result does not depend on the
// schedule.
//
// The expansion centre is the FP64 centre of mass of the node (rounded once
// to float), so the dipole vanishes to rounding.  The opening radius is the
// larger of size/theta + |COM - centre of the node square| and b/theta, b
// the largest distance of the node's particles from the COM (Barnes 1994),
// compared against the plain distance.  The second bound matters because the
// reference's float partition points can leave a particle up to half a node
// outside its node's square (when its Morton key and its float coordinate
// disagree about a midline); a truncated series about such a node diverges
// for targets closer than that particle.  With A_n = C_n / n! (C_n the
// central moments), T_l = A3_iil, U_lm = A4_iilm and W = A4_iijj, the record
// per node is 6 float4 holding the coefficients of the gradient polynomials
// the walk evaluates (g2 = -6 A2 d + 9 T, g3 = -45 A3:dd + 180 U d,
// g4 = -420 A4:ddd) and the constants of the series:
//   0: (comx, comy, opening radius squared, mass)
//   1: (-6 A2xx, -6 A2xy, -6 A2yy, 3 tr A2)
//   2: (9 T_x, 9 T_y, -45 W, 180 U_xy)
//   3: -45 A3: (xxx, xxy, xyy, yyy)
//   4: (180 U_xx, 180 U_yy, -420 A4xxxx, -420 A4xxxy)
//   5: -420 (3 A4xxyy, A4xyyy, A4yyyy, 3 A4xxxy)
// (the factors 3 are g4's binomial weights, stored where a term only
// appears with them)
// The walk evaluates the Taylor expansion of the softened kernel
// (|d|^2 + eps^2)^-1/2 to the order asked for from these; the quadrupole
// reads records 0-1, the octupole 0-3 (without W and U).
#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kMono = 15;                 // monomials x^a y^b, a + b <= 4
constexpr int kPiece = kMono + 1;         // per piece: the monomial sums and the largest offset
constexpr uint32_t kMpThreadMax = 32;     // nodes summed by one thread
constexpr uint32_t kMpWaveMax = 4096;     // nodes summed by one wave; pieces above
constexpr int kMpBlock = 256;

// counters of the multipole pass (mp_cnt)
enum : int { kMpWaveCount = 0, kMpBigCount = 1, kMpPieceCount = 2, kMpSlots = 4 };

__host__ __device__ constexpr int mono(int a, int b) { return (a + b) * (a + b + 1) / 2 + b; }

// particle j as (x, y, m) in FP64: its float position, or with the precision
// form (PREC) its home form, which FP64 holds exactly
struct MonoPt {
    double x, y, m;
};

template <bool PREC>
__device__ __forceinline__ MonoPt mono_pt(const float4 *__restrict__ pxym, const float4 *__restrict__ pho, uint32_t j) {
    const float4 p = pxym[j];
    if (!PREC) return MonoPt{(double)p.x, (double)p.y, (double)p.z};
    const float4 h = pho[j];
    return MonoPt{(double)__float_as_int(h.z) * (double)kHomeCell + (double)h.x,
                  (double)__float_as_int(h.w) * (double)kHomeCell + (double)h.y, (double)p.z};
}

// adds the monomials of particle p about (cx, cy) to S; r2 = largest squared offset
__device__ __forceinline__ void mono_add(const MonoPt p, double cx, double cy, double *__restrict__ S, double &r2) {
    const double m = p.m, x = p.x - cx, y = p.y - cy;
    r2 = fmax(r2, x * x + y * y);
    double xp[5], yp[5];
    xp[0] = m;
    yp[0] = 1.0;
#pragma unroll
    for (int k = 1; k < 5; ++k) {
        xp[k] = xp[k - 1] * x;
        yp[k] = yp[k - 1] * y;
    }
#pragma unroll
    for (int a = 0; a <= 4; ++a)
#pragma unroll
        for (int b = 0; a + b <= 4; ++b) S[mono(a, b)] += xp[a] * yp[b];
}

// fixed-order wave sum (xor butterfly) of S and wave max of r2: every lane
// ends with the totals
__device__ __forceinline__ void wave_sum(double *__restrict__ S, double &r2) {
#pragma unroll
    for (int c = 0; c < kMono; ++c) {
        double v = S[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        S[c] = v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
}

// Half-wave and half-row swaps of a double (v_permlane32_swap /
// v_permlane16_swap on each 32-bit half): with HALF = 32, lanes 32-63 of x
// trade places with lanes 0-31 of y; with HALF = 16 the odd 16-lane rows of
// x trade with the even rows of y.
template <int HALF>
__device__ __forceinline__ void swap_halves(double &x, double &y) {
    const uint64_t xb = (uint64_t)__double_as_longlong(x), yb = (uint64_t)__double_as_longlong(y);
    const uint32_t xl = (uint32_t)xb, xh = (uint32_t)(xb >> 32), yl = (uint32_t)yb, yh = (uint32_t)(yb >> 32);
    uint32_t rxl, ryl, rxh, ryh;
    if (HALF == 32) {
        const auto lo = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
        rxl = lo[0]; ryl = lo[1]; rxh = hi[0]; ryh = hi[1];
    } else {
        const auto lo = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
        rxl = lo[0]; ryl = lo[1]; rxh = hi[0]; ryh = hi[1];
    }
    x = __longlong_as_double((long long)(((uint64_t)rxh << 32) | rxl));
    y = __longlong_as_double((long long)(((uint64_t)ryh << 32) | ryl));
}

// One transposing step over lane bit BIT with W slots: a lane whose bit is
// clear keeps slots [0, W) and the one whose bit is set keeps [W, 2W), each
// adding its partner's copy.  The slot that may hold r2 (the last component)
// takes the max.
template <int W, int BIT>
__device__ __forceinline__ void tstep(double *__restrict__ v, uint32_t lane, uint32_t &cb) {
    const bool hi = (lane >> BIT) & 1u;
    cb += hi ? (uint32_t)W : 0u;
    const bool mx = cb + (uint32_t)W == (uint32_t)(kMono + 1);
#pragma unroll
    for (int i = 0; i < W; ++i) {
        double a = v[i], b = v[W + i];
        if (BIT == 5) {
            swap_halves<32>(a, b);
        } else if (BIT == 4) {
            swap_halves<16>(a, b);
        } else {
            const double keep = hi ? b : a;
            b = __shfl_xor(hi ? a : b, 1 << BIT, 64);
            a = keep;
        }
        v[i] = (i == W - 1 && mx) ? fmax(a, b) : a + b;
    }
}

// The same sums transposed: lanes 4c .. 4c + 3 end with component c (c <
// kMono: the sum of S[c] over the wave; c == kMono: the wave max of r2).
// Each of the first four steps halves the components a lane carries, so a
// wave moves 17 doubles instead of 6 x 16, and the two widest steps are
// lane swaps.  The order of the sums is fixed.
__device__ __forceinline__ double wave_sum_t(const double *__restrict__ S, double r2) {
    const uint32_t lane = threadIdx.x & 63u;
    double v[kMono + 1];
#pragma unroll
    for (int c = 0; c < kMono; ++c) v[c] = S[c];
    v[kMono] = r2;
    uint32_t cb = 0;  // first component of the block this lane keeps
    tstep<8, 5>(v, lane, cb);
    tstep<4, 4>(v, lane, cb);
    tstep<2, 3>(v, lane, cb);
    tstep<1, 2>(v, lane, cb);
    const bool mx = cb == (uint32_t)kMono;
#pragma unroll
    for (int o = 2; o > 0; o >>= 1) {
        const double recv = __shfl_xor(v[0], o, 64);
        v[0] = mx ? fmax(v[0], recv) : v[0] + recv;
    }
    return v[0];
}

// T = S shifted by w: T_ab = sum_{i<=a, j<=b} C(a,i) C(b,j) (-w_x)^(a-i) (-w_y)^(b-j) S_ij
__device__ __forceinline__ void mom_shift_add(const double *__restrict__ S, double wx, double wy, double *__restrict__ T) {
    constexpr double binom[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};
    double px[5], py[5];
    px[0] = py[0] = 1.0;
#pragma unroll
    for (int k = 1; k < 5; ++k) {
        px[k] = px[k - 1] * -wx;
        py[k] = py[k - 1] * -wy;
    }
#pragma unroll
    for (int a = 0; a <= 4; ++a)
#pragma unroll
        for (int b = 0; a + b <= 4; ++b) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i <= a; ++i)
#pragma unroll
                for (int j = 0; j <= b; ++j) t += binom[a][i] * binom[b][j] * px[a - i] * py[b - j] * S[mono(i, j)];
            T[mono(a, b)] += t;
        }
}

// Central moments about the expansion centre from S (about the float centre
// of mass r.comx, r.comy), packed into the node's record.  The centre is the
// exact centre of mass rounded to float, or, with the precision form
// (home != nullptr), to its home form, which is stored in home[k].
__device__ void finish_record(uint32_t k, const NodeRec &r, const float4 g, const double *__restrict__ S, double r2,
                              float inv_theta, float4 *__restrict__ mp, float4 *__restrict__ home) {
    const double cx0 = r.comx, cy0 = r.comy;
    const double M = S[mono(0, 0)];
    const double dx = M > 0.0 ? S[mono(1, 0)] / M : 0.0, dy = M > 0.0 ? S[mono(0, 1)] / M : 0.0;
    double ex = (double)(float)(cx0 + dx), ey = (double)(float)(cy0 + dy);
    if (home) {
        const double cx = cx0 + dx, cy = cy0 + dy;
        const double hx = rint(cx * (double)kHomeInv), hy = rint(cy * (double)kHomeInv);
        const bool fits = fabs(hx) < 2147483520.0 && fabs(hy) < 2147483520.0;
        const float ocx = fits ? (float)(cx - hx * (double)kHomeCell) : (float)cx;
        const float ocy = fits ? (float)(cy - hy * (double)kHomeCell) : (float)cy;
        const int32_t ihx = fits ? (int32_t)hx : 0, ihy = fits ? (int32_t)hy : 0;
        ex = (double)ihx * (double)kHomeCell + (double)ocx;
        ey = (double)ihy * (double)kHomeCell + (double)ocy;
        home[k] = make_float4(ocx, ocy, __int_as_float(ihx), __int_as_float(ihy));
    }
    const float comx = (float)ex, comy = (float)ey;
    double C[kMono];
#pragma unroll
    for (int c = 0; c < kMono; ++c) C[c] = 0.0;
    mom_shift_add(S, ex - cx0, ey - cy0, C);
    const float half = 0.5f * g.z;
    const float ox = comx - (g.x + half), oy = comy - (g.y + half);
    // largest particle offset from the exact COM, bounded through the centre
    const double bmax = sqrt(r2) + sqrt((ex - cx0) * (ex - cx0) + (ey - cy0) * (ey - cy0));
    const float lim = fmaxf(g.z * inv_theta + sqrtf(ox * ox + oy * oy), (float)(bmax * (double)inv_theta) * 1.0000002f);
    const double c2xx = C[mono(2, 0)], c2xy = C[mono(1, 1)], c2yy = C[mono(0, 2)];
    const double c3xxx = C[mono(3, 0)], c3xxy = C[mono(2, 1)], c3xyy = C[mono(1, 2)], c3yyy = C[mono(0, 3)];
    const double c4xxxx = C[mono(4, 0)], c4xxxy = C[mono(3, 1)], c4xxyy = C[mono(2, 2)];
    const double c4xyyy = C[mono(1, 3)], c4yyyy = C[mono(0, 4)];
    const double uxx = c4xxxx + c4xxyy, uxy = c4xxxy + c4xyyy, uyy = c4xxyy + c4yyyy;
    // each moment is stored as the coefficient of the gradient polynomial it
    // enters, so the walk applies no per-node scaling
    float4 *o = mp + 6ull * k;
    o[0] = make_float4(comx, comy, lim * lim, (float)M);
    o[1] = make_float4((float)(-3.0 * c2xx), (float)(-3.0 * c2xy), (float)(-3.0 * c2yy), (float)(1.5 * (c2xx + c2yy)));
    o[2] = make_float4((float)(1.5 * (c3xxx + c3xyy)), (float)(1.5 * (c3xxy + c3yyy)), (float)(-1.875 * (uxx + uyy)),
                       (float)(7.5 * uxy));
    o[3] = make_float4((float)(-7.5 * c3xxx), (float)(-7.5 * c3xxy), (float)(-7.5 * c3xyy), (float)(-7.5 * c3yyy));
    o[4] = make_float4((float)(7.5 * uxx), (float)(7.5 * uyy), (float)(-17.5 * c4xxxx), (float)(-17.5 * c4xxxy));
    o[5] = make_float4((float)(-52.5 * c4xxyy), (float)(-17.5 * c4xyyy), (float)(-17.5 * c4yyyy), (float)(-52.5 * c4xxxy));
}

__global__ void k_mp_reset(uint32_t *__restrict__ cnt) {
    if (threadIdx.x < kMpSlots) cnt[threadIdx.x] = 0u;
}

// mono_add over particles j, j + stride, ... < end, in that order, with four
// loads in flight (the sums keep the order, so the result is the plain loop's)
template <bool PREC>
__device__ __forceinline__ void mono_range(const float4 *__restrict__ pxym, const float4 *__restrict__ pho,
                                           uint32_t j, uint32_t end, uint32_t stride, double cx, double cy,
                                           double *__restrict__ S, double &r2) {
    for (; j < end && end - j > 3 * stride; j += 4 * stride) {
        const MonoPt p0 = mono_pt<PREC>(pxym, pho, j), p1 = mono_pt<PREC>(pxym, pho, j + stride);
        const MonoPt p2 = mono_pt<PREC>(pxym, pho, j + 2 * stride), p3 = mono_pt<PREC>(pxym, pho, j + 3 * stride);
        mono_add(p0, cx, cy, S, r2);
        mono_add(p1, cx, cy, S, r2);
        mono_add(p2, cx, cy, S, r2);
        mono_add(p3, cx, cy, S, r2);
    }
    for (; j < end; j += stride) mono_add(mono_pt<PREC>(pxym, pho, j), cx, cy, S, r2);
}

// Thread per node: small nodes are finished here; the others are listed.
template <bool PREC>
__global__ __launch_bounds__(kMpBlock) void k_mp_small(const NodeRec *__restrict__ nodes,
                                                       const float4 *__restrict__ geom,
                                                       const uint32_t *__restrict__ counters,
                                                       const float4 *__restrict__ pxym,
                                                       const float4 *__restrict__ pho, float inv_theta,
                                                       uint32_t *__restrict__ cnt, uint32_t *__restrict__ wave_list,
                                                       uint32_t wave_cap, uint32_t *__restrict__ big_list,
                                                       uint32_t *__restrict__ big_first, uint32_t big_cap,
                                                       uint32_t *__restrict__ piece_owner, uint32_t piece_cap,
                                                       uint32_t *__restrict__ err, float4 *__restrict__ mp,
                                                       float4 *__restrict__ home) {
    const uint32_t nn = counters[kCntNodeTotal];
    for (uint32_t k = blockIdx.x * kMpBlock + threadIdx.x; k < nn; k += gridDim.x * kMpBlock) {
        const NodeRec r = nodes[k];
        const uint32_t count = r.end > r.start ? r.end - r.start : 0u;
        if (count <= kMpThreadMax) {
            double S[kMono];
#pragma unroll
            for (int c = 0; c < kMono; ++c) S[c] = 0.0;
            const double cx = r.comx, cy = r.comy;
            double r2 = 0.0;
            mono_range<PREC>(pxym, pho, r.start, r.end, 1u, cx, cy, S, r2);
            finish_record(k, r, geom[k], S, r2, inv_theta, mp, home);
        }
        // wave-list entries: one atomic per wave (the list order is free)
        const bool to_wave = count > kMpThreadMax && count <= kMpWaveMax;
        const uint64_t wm = __ballot(to_wave);
        if (wm) {
            const uint32_t lane = threadIdx.x & 63u;
            const uint32_t leader = (uint32_t)__builtin_ctzll(wm);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&cnt[kMpWaveCount], (uint32_t)__popcll(wm));
            base = (uint32_t)__shfl((int)base, (int)leader, 64);
            if (to_wave) {
                const uint32_t w = base + (uint32_t)__popcll(wm & ((1ull << lane) - 1ull));
                if (w < wave_cap) wave_list[w] = k;
                else atomicOr(err, kErrMultipole);
            }
        }
        if (count > kMpWaveMax) {
            const uint32_t pieces = (count + kMpWaveMax - 1) / kMpWaveMax;
            const uint32_t b = atomicAdd(&cnt[kMpBigCount], 1u);
            const uint32_t f = atomicAdd(&cnt[kMpPieceCount], pieces);
            if (b < big_cap && f + pieces <= piece_cap) {
                big_list[b] = k;
                big_first[b] = f;
                for (uint32_t q = 0; q < pieces; ++q) piece_owner[f + q] = b;
            } else {
                atomicOr(err, kErrMultipole);
            }
        }
    }
}

// Wave per listed node (kMpThreadMax < count <= kMpWaveMax).  A wave sums
// up to kMpBatch consecutive list entries one after another and parks each
// total in its own LDS slice; lane t then finishes entry t, so the record's
// arithmetic runs on kMpBatch lanes at once instead of one.
constexpr uint32_t kMpBatch = 8;

template <bool PREC>
__global__ __launch_bounds__(kMpBlock) void k_mp_wave(const NodeRec *__restrict__ nodes,
                                                         const float4 *__restrict__ geom,
                                                         const float4 *__restrict__ pxym,
                                                         const float4 *__restrict__ pho, float inv_theta,
                                                         const uint32_t *__restrict__ cnt,
                                                         const uint32_t *__restrict__ wave_list, uint32_t wave_cap,
                                                         float4 *__restrict__ mp, float4 *__restrict__ home) {
    constexpr uint32_t kWaves = kMpBlock / 64;
    __shared__ double tot[kWaves][kMpBatch][kPiece];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t nw = min(cnt[kMpWaveCount], wave_cap);
    const uint32_t waves = gridDim.x * kWaves;
    for (uint32_t e0 = (blockIdx.x * kWaves + wid) * kMpBatch; e0 < nw; e0 += waves * kMpBatch) {
        const uint32_t nt = min(kMpBatch, nw - e0);
        for (uint32_t t = 0; t < nt; ++t) {
            const NodeRec r = nodes[wave_list[e0 + t]];
            double S[kMono];
#pragma unroll
            for (int c = 0; c < kMono; ++c) S[c] = 0.0;
            const double cx = r.comx, cy = r.comy;
            double r2 = 0.0;
            mono_range<PREC>(pxym, pho, r.start + lane, r.end, 64u, cx, cy, S, r2);
            const double v = wave_sum_t(S, r2);
            if ((lane & 3u) == 0u) tot[wid][t][lane >> 2] = v;
        }
        // the slice is this wave's own: a wave-level fence orders its LDS
        // writes before the reads of other lanes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < nt) {
            const uint32_t k = wave_list[e0 + lane];
            double S[kMono];
#pragma unroll
            for (int c = 0; c < kMono; ++c) S[c] = tot[wid][lane][c];
            finish_record(k, nodes[k], geom[k], S, tot[wid][lane][kMono], inv_theta, mp, home);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Wave per piece of a large node: sums about the node's float centre of mass.
template <bool PREC>
__global__ __launch_bounds__(kMpBlock) void k_mp_pieces(const NodeRec *__restrict__ nodes,
                                                        const float4 *__restrict__ pxym,
                                                        const float4 *__restrict__ pho,
                                                        const uint32_t *__restrict__ cnt,
                                                        const uint32_t *__restrict__ big_list,
                                                        const uint32_t *__restrict__ big_first, uint32_t big_cap,
                                                        const uint32_t *__restrict__ piece_owner, uint32_t piece_cap,
                                                        double *__restrict__ piece_sum) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t np = min(cnt[kMpPieceCount], piece_cap);
    const uint32_t nb = min(cnt[kMpBigCount], big_cap);
    const uint32_t waves = gridDim.x * (kMpBlock / 64);
    for (uint32_t e = blockIdx.x * (kMpBlock / 64) + (threadIdx.x >> 6); e < np; e += waves) {
        const uint32_t b = piece_owner[e];
        if (b >= nb) continue;
        const uint32_t k = big_list[b];
        const NodeRec r = nodes[k];
        const uint32_t lo = r.start + (e - big_first[b]) * kMpWaveMax;
        const uint32_t hi = min(lo + kMpWaveMax, r.end);
        double S[kMono];
#pragma unroll
        for (int c = 0; c < kMono; ++c) S[c] = 0.0;
        const double cx = r.comx, cy = r.comy;
        double r2 = 0.0;
        mono_range<PREC>(pxym, pho, lo + lane, hi, 64u, cx, cy, S, r2);
        // slot kMono: the largest squared offset
        const double v = wave_sum_t(S, r2);
        if ((lane & 3u) == 0u) piece_sum[(size_t)kPiece * e + (lane >> 2)] = v;
    }
}

// Wave per large node: lane l sums pieces l, l + 64, ... in order, then a
// fixed-order wave reduction.
__global__ __launch_bounds__(kMpBlock) void k_mp_big(const NodeRec *__restrict__ nodes,
                                                     const float4 *__restrict__ geom, float inv_theta,
                                                     const uint32_t *__restrict__ cnt,
                                                     const uint32_t *__restrict__ big_list,
                                                     const uint32_t *__restrict__ big_first, uint32_t big_cap,
                                                     uint32_t piece_cap, const double *__restrict__ piece_sum,
                                                     float4 *__restrict__ mp, float4 *__restrict__ home) {
    const uint32_t nb = min(cnt[kMpBigCount], big_cap);
    const uint32_t np = cnt[kMpPieceCount];
    if (np > piece_cap) return;  // the error flag is already raised
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (kMpBlock / 64);
    for (uint32_t b = blockIdx.x * (kMpBlock / 64) + (threadIdx.x >> 6); b < nb; b += waves) {
        const uint32_t k = big_list[b];
        const NodeRec r = nodes[k];
        const uint32_t pieces = (r.end - r.start + kMpWaveMax - 1) / kMpWaveMax;
        double S[kMono];
#pragma unroll
        for (int c = 0; c < kMono; ++c) S[c] = 0.0;
        double r2 = 0.0;
        for (uint32_t q = lane; q < pieces; q += 64) {
            const double *ps = piece_sum + (size_t)kPiece * (big_first[b] + q);
#pragma unroll
            for (int c = 0; c < kMono; ++c) S[c] += ps[c];
            r2 = fmax(r2, ps[kMono]);
        }
        wave_sum(S, r2);
        if (lane == 0) finish_record(k, r, geom[k], S, r2, inv_theta, mp, home);
    }
}

}  // namespace

// Device buffers of the accuracy modes, allocated on first use (never inside
// a graph capture: ge_step calls this before capturing).  A node above
// kMpWaveMax particles lies on a tree level holding at most n / kMpWaveMax
// such nodes, so the pieces are bounded by 2 n / kMpWaveMax per level.
int ensure_multipole_buffers(ge_ctx *c) {
    Buffers &b = c->b;
    if (b.node_mp) return GE_OK;
    b.mp_wave_cap = b.node_cap;
    b.mp_big_cap = (uint32_t)(((uint64_t)c->cap / kMpWaveMax + 1) * 64 + 1024);
    b.mp_piece_cap = (uint32_t)(((uint64_t)c->cap / kMpWaveMax + 1) * 128 + 1024);
    GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&b.node_mp), sizeof(float4) * 6ull * b.node_cap));
    if (c->precise)
        GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&b.node_home), sizeof(float4) * b.node_cap));
    GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&b.mp_cnt), sizeof(uint32_t) * kMpSlots));
    GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&b.mp_wave_list), sizeof(uint32_t) * b.mp_wave_cap));
    GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&b.mp_big_list), sizeof(uint32_t) * 2ull * b.mp_big_cap));
    GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&b.mp_piece_owner), sizeof(uint32_t) * b.mp_piece_cap));
    GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&b.mp_piece_sum), sizeof(double) * kPiece * b.mp_piece_cap));
    return GE_OK;
}

void launch_multipoles(ge_ctx *c, float theta) {
    Buffers &b = c->b;
    const float inv_theta = 1.0f / theta;
    uint32_t *big_first = b.mp_big_list + b.mp_big_cap;
    hipLaunchKernelGGL(k_mp_reset, dim3(1), dim3(64), 0, c->stream, b.mp_cnt);
    uint32_t nb = blocks_for(b.node_cap, kMpBlock);
    if (nb > 8192) nb = 8192;
    const float4 *pho = c->precise ? b.pho : nullptr;
    float4 *home = c->precise ? b.node_home : nullptr;
    if (c->precise) {
        hipLaunchKernelGGL(k_mp_small<true>, dim3(nb), dim3(kMpBlock), 0, c->stream, b.nodes, b.node_geom, b.counters,
                           b.pxym, pho, inv_theta, b.mp_cnt, b.mp_wave_list, b.mp_wave_cap, b.mp_big_list, big_first,
                           b.mp_big_cap, b.mp_piece_owner, b.mp_piece_cap, b.counters + kCntError, b.node_mp, home);
        hipLaunchKernelGGL(k_mp_wave<true>, dim3(2048), dim3(kMpBlock), 0, c->stream, b.nodes, b.node_geom, b.pxym, pho,
                           inv_theta, b.mp_cnt, b.mp_wave_list, b.mp_wave_cap, b.node_mp, home);
        hipLaunchKernelGGL(k_mp_pieces<true>, dim3(1024), dim3(kMpBlock), 0, c->stream, b.nodes, b.pxym, pho, b.mp_cnt,
                           b.mp_big_list, big_first, b.mp_big_cap, b.mp_piece_owner, b.mp_piece_cap, b.mp_piece_sum);
    } else {
        hipLaunchKernelGGL(k_mp_small<false>, dim3(nb), dim3(kMpBlock), 0, c->stream, b.nodes, b.node_geom, b.counters,
                           b.pxym, pho, inv_theta, b.mp_cnt, b.mp_wave_list, b.mp_wave_cap, b.mp_big_list, big_first,
                           b.mp_big_cap, b.mp_piece_owner, b.mp_piece_cap, b.counters + kCntError, b.node_mp, home);
        hipLaunchKernelGGL(k_mp_wave<false>, dim3(2048), dim3(kMpBlock), 0, c->stream, b.nodes, b.node_geom, b.pxym,
                           pho, inv_theta, b.mp_cnt, b.mp_wave_list, b.mp_wave_cap, b.node_mp, home);
        hipLaunchKernelGGL(k_mp_pieces<false>, dim3(1024), dim3(kMpBlock), 0, c->stream, b.nodes, b.pxym, pho,
                           b.mp_cnt, b.mp_big_list, big_first, b.mp_big_cap, b.mp_piece_owner, b.mp_piece_cap,
                           b.mp_piece_sum);
    }
    hipLaunchKernelGGL(k_mp_big, dim3(256), dim3(kMpBlock), 0, c->stream, b.nodes, b.node_geom, inv_theta, b.mp_cnt,
                       b.mp_big_list, big_first, b.mp_big_cap, b.mp_piece_cap, b.mp_piece_sum, b.node_mp, home);
}

}  // namespace ge
