// Drives the 3D C++ facade the way mode3D() drives the reference's 3D step
// (GE/src/globalLogic.cpp:1580-1719) with SPH3D on: stage by stage with the
// reference's call names, and fused (ge3_step).  Both run the same
// arithmetic, so the program exits 0 only if their states agree bit for bit.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "ge_facade3d.hpp"

static ge_gpu::Particles3D ball(uint32_t n) {
    ge_gpu::Particles3D p;
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    const float radius = std::cbrt((float)n / 0.3f / 4.18879f);
    for (uint32_t i = 0; i < n; ++i) {
        float x, y, z;
        do {
            x = 2.0f * u(rng) - 1.0f;
            y = 2.0f * u(rng) - 1.0f;
            z = 2.0f * u(rng) - 1.0f;
        } while (x * x + y * y + z * z > 1.0f);
        p.x.push_back(radius * x);
        p.y.push_back(radius * y);
        p.z.push_back(radius * z);
        p.vx.push_back(u(rng) - 0.5f);
        p.vy.push_back(u(rng) - 0.5f);
        p.vz.push_back(u(rng) - 0.5f);
        p.mass.push_back(1.0f);
        p.flags.push_back(GE_FLAG_SPH);
        p.id.push_back(i);
        p.sph_mass.push_back(0.9f + 0.2f * u(rng));
        p.rest_dens.push_back(0.134f);
        p.stiff.push_back(1.0f);
        p.visc.push_back(1.0f);
        p.cohesion.push_back(1.0f);
    }
    return p;
}

int main(int argc, char **argv) {
    const uint32_t n = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 20000;
    const int frames = 3;
    ge3_step_params myVar;
    ge3_default_params(&myVar);
    myVar.G = 1.0f;
    myVar.theta = 0.7f;
    myVar.softening = 1.0f;
    myVar.dt = myVar.dt_start = 0.002f;
    myVar.direct_below = 0;
    myVar.sph = 1;
    try {
        ge_gpu::Particles3D init = ball(n);
        ge_gpu::Scene3D staged(n, true), fused(n, true);
        staged.upload(init);
        fused.upload(init);
        ge_gpu::Morton3D morton;
        ge_gpu::Physics3D physics;
        ge_gpu::NeighborSearch3DV2AVX2 grid;
        ge_gpu::SPH3D sph;
        size_t nodes = 0;
        for (int f = 0; f < frames; ++f) {
            physics.integrateStart3D(staged, myVar);
            const auto bb = ge_gpu::boundingBox3D(staged);
            morton.computeMortonKeys3D(staged, bb);
            morton.sortParticlesByMortonKey3D(staged);
            ge_gpu::Octree root(staged, bb);
            nodes = root.nodes().size();
            grid.newGridAVX2(staged, myVar);
            physics.calculateForceFromGrid3D(staged, myVar);
            sph.pcisphSolver(staged, myVar);
            physics.integrateEnd3D(staged, myVar);
        }
        ge_gpu::mode3D(fused, myVar, frames);
        ge_gpu::Particles3D a, b;
        for (ge_gpu::Particles3D *p : {&a, &b}) {
            for (auto *v : {&p->x, &p->y, &p->z, &p->vx, &p->vy, &p->vz, &p->press, &p->pfx}) v->resize(1);
            p->id.resize(1);
        }
        staged.download(a);
        fused.download(b);
        const bool same = a.x == b.x && a.y == b.y && a.z == b.z && a.vx == b.vx && a.vy == b.vy && a.vz == b.vz &&
                          a.press == b.press && a.pfx == b.pfx && a.id == b.id;
        std::printf("facade3d: n=%u frames=%d nodes=%zu staged==fused: %s\n", n, frames, nodes, same ? "yes" : "NO");
        return same ? 0 : 1;
    } catch (const ge_gpu::Error &e) {
        std::printf("facade3d: error %d: %s\n", e.code, e.what());
        return 2;
    }
}
