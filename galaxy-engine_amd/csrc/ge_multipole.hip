// ge_multipole.hip -- node quadrupole moments and opening radii for the
// accuracy mode (ge_step_params.accuracy = 1).  This mode has no counterpart
// in the reference, whose walk uses monopoles only (physics.cpp:250-259).
//
// Quadrupole of a node about its centre of mass c, for the 1/r kernel of the
// reference restricted to the plane:
//   Q = sum_i m_i (3 s_i s_i^T - |s_i|^2 I),  s_i = x_i - c.
// Over the Morton-sorted particles every node is a contiguous range, so its
// raw moments (sum m, m x, m y, m x^2, m x y, m y^2, taken about the box
// centre) are differences of FP64 prefix sums, and the central moments
// follow in O(1) per node.  The opening radius is size/theta + |c - centre of
// the node square| (Barnes 1994), compared against the plain distance.
#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kMomThreads = 1024;
constexpr int kMomItems = kChunk / kMomThreads;  // one workgroup per reduction chunk

struct Mom {
    double v[6];
};

// raw moments are taken about the centre of the particle's system box
__device__ __forceinline__ void box_centre(const float *__restrict__ sys_bb, uint32_t sy, double &ox, double &oy) {
    const float *bb = sys_bb + 8ull * sy;
    ox = (double)bb[0] + 0.5 * (double)bb[2];
    oy = (double)bb[1] + 0.5 * (double)bb[2];
}

__device__ __forceinline__ Mom mom_of(const float4 *__restrict__ pxym, double ox, double oy, uint32_t i, uint32_t hi) {
    Mom r;
    if (i < hi) {
        const float4 p = pxym[i];
        const double m = p.z, x = (double)p.x - ox, y = (double)p.y - oy;
        r.v[0] = m; r.v[1] = m * x; r.v[2] = m * y; r.v[3] = m * x * x; r.v[4] = m * x * y; r.v[5] = m * y * y;
    } else {
        for (int c = 0; c < 6; ++c) r.v[c] = 0.0;
    }
    return r;
}

// Exclusive workgroup scan of a 6-vector with a fixed addition tree: the
// prefix of an item depends only on the items before it.  tot = group total.
__device__ Mom block_scan6(Mom v, Mom &tot) {
    __shared__ double sh[kMomThreads / 64 + 1][6];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    Mom incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        for (int c = 0; c < 6; ++c) {
            const double u = __shfl_up(incl.v[c], d, 64);
            if (lane >= (uint32_t)d) incl.v[c] += u;
        }
    }
    Mom excl;
    for (int c = 0; c < 6; ++c) {
        const double u = __shfl_up(incl.v[c], 1, 64);
        excl.v[c] = lane == 0 ? 0.0 : u;
    }
    if (lane == 63)
        for (int c = 0; c < 6; ++c) sh[w][c] = incl.v[c];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        double run = 0.0;
        for (int k = 0; k < kMomThreads / 64; ++k) {
            const double t = sh[k][c];
            sh[k][c] = run;
            run += t;
        }
        sh[kMomThreads / 64][c] = run;
    }
    __syncthreads();
    Mom ex;
    for (int c = 0; c < 6; ++c) {
        ex.v[c] = w == 0 ? excl.v[c] : sh[w][c] + excl.v[c];
        tot.v[c] = sh[kMomThreads / 64][c];
    }
    __syncthreads();
    return ex;
}

// Per chunk (chunks never span systems and start at the system's first
// particle + multiples of kChunk): thread t owns items lo + t*kMomItems ...
__global__ __launch_bounds__(kMomThreads) void k_mom_sums(const float4 *__restrict__ pxym,
                                                          const uint32_t *__restrict__ chunk_lo,
                                                          const uint32_t *__restrict__ sys_of,
                                                          const float *__restrict__ sys_bb, double *__restrict__ part) {
    const uint32_t lo = chunk_lo[blockIdx.x], hi = chunk_lo[blockIdx.x + 1];
    Mom tot;
    for (int c = 0; c < 6; ++c) tot.v[c] = 0.0;
    if (lo < hi) {
        double ox, oy;
        box_centre(sys_bb, sys_of[lo], ox, oy);
        const uint32_t base = lo + threadIdx.x * kMomItems;
        Mom s;
        for (int c = 0; c < 6; ++c) s.v[c] = 0.0;
        for (int k = 0; k < kMomItems; ++k) {
            const Mom m = mom_of(pxym, ox, oy, base + k, hi);
            for (int c = 0; c < 6; ++c) s.v[c] += m.v[c];
        }
        block_scan6(s, tot);
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) part[6 * blockIdx.x + c] = tot.v[c];
}

// chunk offsets within each system, chunks in order (one thread per system)
__global__ __launch_bounds__(kBlock) void k_mom_parts(double *__restrict__ part, const uint32_t *__restrict__ chunk_first,
                                                      uint32_t nsys) {
    for (uint32_t sy = blockIdx.x * kBlock + threadIdx.x; sy < nsys; sy += gridDim.x * kBlock) {
        double run[6] = {0, 0, 0, 0, 0, 0};
        for (uint32_t k = chunk_first[sy]; k < chunk_first[sy + 1]; ++k) {
            for (int c = 0; c < 6; ++c) {
                const double t = part[6 * k + c];
                part[6 * k + c] = run[c];
                run[c] += t;
            }
        }
    }
}

// mom[6 (i + s) + c] = sum over the particles of system s before slot i,
// for i = off_s .. off_{s+1} (one extra entry per system)
__global__ __launch_bounds__(kMomThreads) void k_mom_apply(const float4 *__restrict__ pxym,
                                                           const uint32_t *__restrict__ chunk_lo,
                                                           const uint32_t *__restrict__ sys_of,
                                                           const uint32_t *__restrict__ sys_off,
                                                           const float *__restrict__ sys_bb,
                                                           const double *__restrict__ part,
                                                           double *__restrict__ mom) {
    const uint32_t lo = chunk_lo[blockIdx.x], hi = chunk_lo[blockIdx.x + 1];
    if (lo >= hi) return;
    const uint32_t sy = sys_of[lo];
    double ox, oy;
    box_centre(sys_bb, sy, ox, oy);
    const uint32_t base = lo + threadIdx.x * kMomItems;
    Mom item[kMomItems];
    Mom s;
    for (int c = 0; c < 6; ++c) s.v[c] = 0.0;
    for (int k = 0; k < kMomItems; ++k) {
        item[k] = mom_of(pxym, ox, oy, base + k, hi);
        for (int c = 0; c < 6; ++c) s.v[c] += item[k].v[c];
    }
    Mom tot;
    Mom run = block_scan6(s, tot);
    for (int c = 0; c < 6; ++c) run.v[c] += part[6 * blockIdx.x + c];
    const bool last_chunk = hi == sys_off[sy + 1];
    for (int k = 0; k < kMomItems; ++k) {
        const uint32_t i = base + k;
        if (i < hi || (last_chunk && i == hi))
            for (int c = 0; c < 6; ++c) mom[6ull * (i + sy) + c] = run.v[c];
        for (int c = 0; c < 6; ++c) run.v[c] += item[k].v[c];
    }
}

__global__ __launch_bounds__(kBlock) void k_node_quad(const NodeRec *__restrict__ nodes,
                                                      const float4 *__restrict__ geom,
                                                      const uint32_t *__restrict__ counters,
                                                      const double *__restrict__ mom,
                                                      const uint32_t *__restrict__ sys_of,
                                                      const float *__restrict__ sys_bb,
                                                      float inv_theta, float4 *__restrict__ quad) {
    const uint32_t nn = counters[kCntNodeTotal];
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < nn; k += gridDim.x * kBlock) {
        const NodeRec r = nodes[k];
        const uint32_t sy = sys_of[r.start];
        double ox, oy;
        box_centre(sys_bb, sy, ox, oy);
        const float4 g = geom[k];
        const double *a = mom + 6ull * (r.start + sy);
        const double *b = mom + 6ull * (r.end + sy);
        const double M = b[0] - a[0], Sx = b[1] - a[1], Sy = b[2] - a[2];
        const double Sxx = b[3] - a[3], Sxy = b[4] - a[4], Syy = b[5] - a[5];
        const double cx = (double)r.comx - ox, cy = (double)r.comy - oy;
        const double Cxx = Sxx - 2.0 * cx * Sx + cx * cx * M;
        const double Cyy = Syy - 2.0 * cy * Sy + cy * cy * M;
        const double Cxy = Sxy - cx * Sy - cy * Sx + cx * cy * M;
        const float half = 0.5f * g.z;
        const float ddx = r.comx - (g.x + half), ddy = r.comy - (g.y + half);
        const float lim = g.z * inv_theta + sqrtf(ddx * ddx + ddy * ddy);
        quad[k] = make_float4((float)(2.0 * Cxx - Cyy), (float)(3.0 * Cxy), (float)(2.0 * Cyy - Cxx), lim);
    }
}

}  // namespace

void launch_quad_moments(ge_ctx *c, float theta) {
    hipLaunchKernelGGL(k_mom_sums, dim3(c->nchunks), dim3(kMomThreads), 0, c->stream, c->b.pxym, c->b.chunk_lo,
                       c->b.sys_of, c->b.sys_bb, c->b.moment_part);
    hipLaunchKernelGGL(k_mom_parts, dim3(blocks_for(c->nsys, kBlock)), dim3(kBlock), 0, c->stream, c->b.moment_part,
                       c->b.chunk_first, c->nsys);
    hipLaunchKernelGGL(k_mom_apply, dim3(c->nchunks), dim3(kMomThreads), 0, c->stream, c->b.pxym, c->b.chunk_lo,
                       c->b.sys_of, c->b.sys_off, c->b.sys_bb, c->b.moment_part, c->b.moments);
    uint32_t nb = blocks_for(c->b.node_cap, kBlock);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_node_quad, dim3(nb), dim3(kBlock), 0, c->stream, c->b.nodes, c->b.node_geom, c->b.counters,
                       c->b.moments, c->b.sys_of, c->b.sys_bb, 1.0f / theta, c->b.node_quad);
}

}  // namespace ge
