// ge_diag.hip -- FP64 diagnostics: conserved-quantity reductions and an FP64
// direct-sum force reference on the device.  The reference shows none of
// these (its UI reports particle counts and FPS only, GE/src/UI/UI.cpp);
// they are the measurement side of the accuracy form: FP64 is used here and
// nowhere on the step itself.
#include <atomic>
#include <mutex>

#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kDiagThreads = 256;

// Fixed-order FP64 block sum: the result depends only on the inputs.
template <int NV>
__device__ void block_sum(double (&v)[NV], double *__restrict__ out) {
    __shared__ double sh[kDiagThreads / 64][NV];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o, 64);
    }
    if (lane == 0)
        for (int c = 0; c < NV; ++c) sh[w][c] = v[c];
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int k = 0; k < kDiagThreads / 64; ++k) s += sh[k][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// Per block over slots [lo, hi): sum m, m vx, m vy, m (x vy - y vx),
// m |v|^2 / 2, m x, m y.
__global__ __launch_bounds__(kDiagThreads) void k_moments(State s, uint32_t lo, uint32_t hi,
                                                          double *__restrict__ part) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = lo + blockIdx.x * kDiagThreads + threadIdx.x; i < hi; i += gridDim.x * kDiagThreads) {
        const double m = s.m[i], x = s.x[i], y = s.y[i], vx = s.vx[i], vy = s.vy[i];
        v[0] += m;
        v[1] += m * vx;
        v[2] += m * vy;
        v[3] += m * (x * vy - y * vx);
        v[4] += 0.5 * m * (vx * vx + vy * vy);
        v[5] += m * x;
        v[6] += m * y;
    }
    block_sum<7>(v, part + 7ull * blockIdx.x);
}

// Per block of targets i in [lo, hi): sum over j != i in [lo, hi) of
// m_i m_j / sqrt(r^2 + eps^2), sources staged through LDS.
__global__ __launch_bounds__(kDiagThreads) void k_potential(const float4 *__restrict__ pxym, uint32_t lo, uint32_t hi,
                                                            double eps2, double *__restrict__ part) {
    __shared__ double tx[kDiagThreads], ty[kDiagThreads], tm[kDiagThreads];
    const uint32_t i = lo + blockIdx.x * kDiagThreads + threadIdx.x;
    const bool live = i < hi;
    const float4 p = live ? pxym[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double xi = p.x, yi = p.y;
    double acc = 0.0;
    for (uint32_t base = lo; base < hi; base += kDiagThreads) {
        const uint32_t j = base + threadIdx.x;
        if (j < hi) {
            const float4 q = pxym[j];
            tx[threadIdx.x] = q.x; ty[threadIdx.x] = q.y; tm[threadIdx.x] = q.z;
        } else {
            tx[threadIdx.x] = 0.0; ty[threadIdx.x] = 0.0; tm[threadIdx.x] = 0.0;
        }
        __syncthreads();
        const uint32_t lim = min((uint32_t)kDiagThreads, hi - base);
        for (uint32_t k = 0; k < lim; ++k) {
            if (base + k == i) continue;
            const double dx = tx[k] - xi, dy = ty[k] - yi;
            acc += tm[k] / sqrt(dx * dx + dy * dy + eps2);
        }
        __syncthreads();
    }
    double v[1] = {live ? (double)p.z * acc : 0.0};
    block_sum<1>(v, part + blockIdx.x);
}

// One workgroup per target slot t: FP64 sum over the target's system of
// G m_j d / (|d|^2 + eps^2)^1.5 (the softened kernel of the tree walk),
// self excluded.
__global__ __launch_bounds__(kDiagThreads) void k_direct_fp64(const float4 *__restrict__ pxym,
                                                              const uint32_t *__restrict__ targets,
                                                              const uint32_t *__restrict__ sys_of,
                                                              const uint32_t *__restrict__ sys_off, double G,
                                                              double eps2, double *__restrict__ out) {
    const uint32_t t = targets[blockIdx.x];
    const uint32_t sy = sys_of[t];
    const uint32_t lo = sys_off[sy], hi = sys_off[sy + 1];
    const float4 p = pxym[t];
    const double xi = p.x, yi = p.y;
    double v[2] = {0.0, 0.0};
    for (uint32_t j = lo + threadIdx.x; j < hi; j += kDiagThreads) {
        if (j == t) continue;
        const float4 q = pxym[j];
        const double dx = (double)q.x - xi, dy = (double)q.y - yi;
        const double r2 = dx * dx + dy * dy + eps2;
        const double inv = 1.0 / (r2 * sqrt(r2));
        v[0] += q.z * dx * inv;
        v[1] += q.z * dy * inv;
    }
    v[0] *= G;
    v[1] *= G;
    block_sum<2>(v, out + 2ull * blockIdx.x);
}

// pxym of the current state (the packed stream is only refreshed by the sort)
__global__ __launch_bounds__(kBlock) void k_pack_pxym(State s, uint32_t n, float4 *__restrict__ pxym) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) pxym[i] = make_float4(s.x[i], s.y[i], s.m[i], 0.0f);
}

// Final sums of the block partials on the device: thread k < 7 adds moment k
// of blocks 0, 1, ..., nb-1 in that order, thread 7 the nbp potential
// partials, each from 0.0, which is the order and rounding of the host loop
// in ge_diagnostics, so both entry points return the same bits.
__global__ __launch_bounds__(64) void k_diag_final(const double *__restrict__ part, uint32_t nb, uint32_t nbp,
                                                    double *__restrict__ sums) {
    const uint32_t k = threadIdx.x;
    if (k < 7) {
        double v = 0.0;
        for (uint32_t b = 0; b < nb; ++b) v += part[7ull * b + k];
        sums[k] = v;
    } else if (k == 7) {
        double v = 0.0;
        for (uint32_t b = 0; b < nbp; ++b) v += part[7ull * nb + b];
        sums[7] = v;
    }
}

// Scratch of ge_diagnostics_noalloc: moment partials of at most
// kNoAllocMomentBlocks blocks, potential partials of at most
// kNoAllocPotentialBlocks blocks, then the 8 final sums.  188 KiB per device,
// held until the process ends.
constexpr uint32_t kNoAllocMomentBlocks = 1024;
constexpr uint32_t kNoAllocPotentialBlocks = 16384;  // potential over at most 4,194,304 particles
constexpr size_t kNoAllocDoubles = 7ull * kNoAllocMomentBlocks + kNoAllocPotentialBlocks + 8;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

}  // namespace

}  // namespace ge

using namespace ge;

namespace {

int diag_fail(ge_ctx *c, int code, const char *msg) {
    c->err = msg;
    return code;
}

// Device address of `scratch` per device.  In host code the array's name is
// only a stand-in for the symbol, so kernels and copies get this address,
// looked up once per device by ge_noalloc_setup with that device current.
constexpr int kMaxDevices = 64;
std::atomic<double *> g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

}  // namespace

namespace ge {

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return GE_E_NODEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device].load()) return GE_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return GE_E_HIP;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return GE_E_HIP;
    g_scratch[device].store(static_cast<double *>(ptr));
    return GE_OK;
}

}  // namespace ge

namespace {

// device scratch for partial sums; allocated on first use, kept by the context
int diag_scratch(ge_ctx *c, size_t doubles, double *&out) {
    if (c->diag_cap < doubles) {
        if (c->diag) (void)hipFree(c->diag);
        c->diag = nullptr;
        c->diag_cap = 0;
        GE_HIP_CHECK(c, hipMalloc(reinterpret_cast<void **>(&c->diag), sizeof(double) * doubles));
        c->diag_cap = doubles;
    }
    out = c->diag;
    return GE_OK;
}

}  // namespace

extern "C" {

int ge_diagnostics(ge_ctx *c, uint32_t sys, const ge_step_params *p, int32_t with_potential, ge_diag *out) {
    if (!c || !out || (with_potential && !p)) return GE_E_INVALID;
    if (c->n == 0 || sys >= c->nsys) return diag_fail(c, GE_E_INVALID, "no such system");
    (void)hipSetDevice(c->device);
    const uint32_t lo = c->host_off[sys], hi = c->host_off[sys + 1];
    const uint32_t cnt = hi - lo;
    const State &s = c->b.st[c->b.cur];
    uint32_t nb = blocks_for(cnt, kDiagThreads);
    if (nb > 1024) nb = 1024;
    const uint32_t nbp = with_potential ? blocks_for(cnt, kDiagThreads) : 0u;
    double *part = nullptr;
    if (int r = diag_scratch(c, 7ull * nb + nbp + 8, part)) return r;
    hipLaunchKernelGGL(k_moments, dim3(nb), dim3(kDiagThreads), 0, c->stream, s, lo, hi, part);
    if (with_potential) {
        hipLaunchKernelGGL(k_pack_pxym, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s, c->n,
                           c->b.pxym);
        const double eps2 = (double)p->softening * (double)p->softening;
        hipLaunchKernelGGL(k_potential, dim3(nbp), dim3(kDiagThreads), 0, c->stream, c->b.pxym, lo, hi, eps2,
                           part + 7ull * nb);
    }
    std::vector<double> h(7ull * nb + nbp);
    GE_HIP_CHECK(c, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = 0; b < nb; ++b)
        for (int k = 0; k < 7; ++k) v[k] += h[7ull * b + k];
    double pot = 0.0;
    for (uint32_t b = 0; b < nbp; ++b) pot += h[7ull * nb + b];
    out->mass = v[0];
    out->px = v[1];
    out->py = v[2];
    out->lz = v[3];
    out->kinetic = v[4];
    out->comx = v[0] > 0 ? v[5] / v[0] : 0.0;
    out->comy = v[0] > 0 ? v[6] / v[0] : 0.0;
    out->potential = with_potential ? -0.5 * (double)p->G * pot : 0.0;
    return GE_OK;
}

int ge_diagnostics_noalloc(ge_ctx *c, uint32_t sys, const ge_step_params *p, int32_t with_potential,
                           ge_diag *out) {
    if (!c || !out || (with_potential && !p)) return GE_E_INVALID;
    if (c->n == 0 || sys >= c->nsys) return diag_fail(c, GE_E_INVALID, "no such system");
    double *part = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device].load() : nullptr;
    if (!part) return diag_fail(c, GE_E_HIP, "no scratch address for this device");
    const uint32_t lo = c->host_off[sys], hi = c->host_off[sys + 1];
    const uint32_t cnt = hi - lo;
    uint32_t nb = blocks_for(cnt, kDiagThreads);
    if (nb > kNoAllocMomentBlocks) nb = kNoAllocMomentBlocks;
    const uint32_t nbp = with_potential ? blocks_for(cnt, kDiagThreads) : 0u;
    if (nbp > kNoAllocPotentialBlocks)
        return diag_fail(c, GE_E_CAPACITY, "with_potential: the system has more than 4194304 particles "
                                           "(the fixed scratch of ge_diagnostics_noalloc)");
    (void)hipSetDevice(c->device);
    double *sums = part + 7ull * kNoAllocMomentBlocks + kNoAllocPotentialBlocks;
    const State &s = c->b.st[c->b.cur];
    hipLaunchKernelGGL(k_moments, dim3(nb), dim3(kDiagThreads), 0, c->stream, s, lo, hi, part);
    if (with_potential) {
        hipLaunchKernelGGL(k_pack_pxym, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s, c->n,
                           c->b.pxym);
        const double eps2 = (double)p->softening * (double)p->softening;
        hipLaunchKernelGGL(k_potential, dim3(nbp), dim3(kDiagThreads), 0, c->stream, c->b.pxym, lo, hi, eps2,
                           part + 7ull * nb);
    }
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, c->stream, part, nb, nbp, sums);
    double v[8];
    GE_HIP_CHECK(c, hipMemcpyAsync(v, sums, sizeof v, hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    out->mass = v[0];
    out->px = v[1];
    out->py = v[2];
    out->lz = v[3];
    out->kinetic = v[4];
    out->comx = v[0] > 0 ? v[5] / v[0] : 0.0;
    out->comy = v[0] > 0 ? v[6] / v[0] : 0.0;
    out->potential = with_potential ? -0.5 * (double)p->G * v[7] : 0.0;
    return GE_OK;
}

int ge_direct_fp64(ge_ctx *c, uint32_t ntargets, const uint32_t *targets, const ge_step_params *p, double *ax,
                   double *ay) {
    if (!c || !p || (ntargets && (!targets || !ax || !ay))) return GE_E_INVALID;
    if (ntargets == 0) return GE_OK;
    for (uint32_t k = 0; k < ntargets; ++k)
        if (targets[k] >= c->n) return diag_fail(c, GE_E_INVALID, "target slot out of range");
    (void)hipSetDevice(c->device);
    double *buf = nullptr;
    // targets (as uint32) then 2 doubles per target
    const size_t tdoubles = (ntargets + 1) / 2;
    if (int r = diag_scratch(c, tdoubles + 2ull * ntargets, buf)) return r;
    uint32_t *dt = reinterpret_cast<uint32_t *>(buf);
    double *res = buf + tdoubles;
    GE_HIP_CHECK(c, hipMemcpyAsync(dt, targets, sizeof(uint32_t) * ntargets, hipMemcpyHostToDevice, c->stream));
    const State &s = c->b.st[c->b.cur];
    hipLaunchKernelGGL(k_pack_pxym, dim3(blocks_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, s, c->n, c->b.pxym);
    const double eps2 = (double)p->softening * (double)p->softening;
    hipLaunchKernelGGL(k_direct_fp64, dim3(ntargets), dim3(kDiagThreads), 0, c->stream, c->b.pxym, dt, c->b.sys_of,
                       c->b.sys_off, (double)p->G, eps2, res);
    std::vector<double> h(2ull * ntargets);
    GE_HIP_CHECK(c, hipMemcpyAsync(h.data(), res, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    GE_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < ntargets; ++k) {
        ax[k] = h[2ull * k];
        ay[k] = h[2ull * k + 1];
    }
    return GE_OK;
}

}  // extern "C"
