// Drives the C++ facade the way updateScene() drives the reference's 2D step
// (GE/src/globalLogic.cpp:1155-1372): stage by stage with the reference's
// call names, and fused (ge_step).  Both orders run the same arithmetic, so
// the program exits 0 only if their states agree bit for bit.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "ge_facade.hpp"

static ge_gpu::Particles disk(uint32_t n) {
    ge_gpu::Particles p;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    for (uint32_t i = 0; i < n; ++i) {
        const float r = 5.0f + 200.0f * std::sqrt(u(rng)), a = 6.2831853f * u(rng);
        const float v = std::sqrt(2000.0f / r);
        p.x.push_back(r * std::cos(a));
        p.y.push_back(r * std::sin(a));
        p.vx.push_back(-v * std::sin(a));
        p.vy.push_back(v * std::cos(a));
        p.mass.push_back(0.5f + u(rng));
        p.id.push_back(i);
    }
    return p;
}

int main(int argc, char **argv) {
    const uint32_t n = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 20000;
    const int frames = 3;
    ge_step_params myVar;
    ge_default_params(&myVar);
    myVar.G = 1.0f;
    myVar.theta = 0.8f;
    myVar.softening = 1.0f;
    myVar.dt = myVar.dt_start = 0.01f;
    myVar.direct_below = 0;
    try {
        ge_gpu::Particles init = disk(n);
        ge_gpu::Scene staged(n), fused(n);
        staged.upload(init);
        fused.upload(init);
        ge_gpu::Morton morton;
        ge_gpu::Physics physics;
        size_t nodes = 0;
        for (int f = 0; f < frames; ++f) {
            physics.integrateStart(staged, myVar);
            const auto bb = ge_gpu::boundingBox(staged);
            morton.computeMortonKeys(staged, bb);
            morton.sortParticlesByMortonKey(staged);
            ge_gpu::Quadtree root(staged, bb);
            nodes = root.nodes().size();
            physics.calculateForceFromGrid(staged, myVar);
            physics.integrateEnd(staged, myVar);
        }
        ge_gpu::updateScene(fused, myVar, frames);
        ge_gpu::Particles a, b;
        for (ge_gpu::Particles *p : {&a, &b}) {
            p->x.resize(1); p->y.resize(1); p->vx.resize(1); p->vy.resize(1); p->id.resize(1);
        }
        staged.download(a);
        fused.download(b);
        const bool same = a.x == b.x && a.y == b.y && a.vx == b.vx && a.vy == b.vy && a.id == b.id;
        std::printf("facade: n=%u frames=%d nodes=%zu staged==fused: %s\n", n, frames, nodes, same ? "yes" : "NO");
        return same ? 0 : 1;
    } catch (const ge_gpu::Error &e) {
        std::printf("facade: error %d: %s\n", e.code, e.what());
        return 2;
    }
}
