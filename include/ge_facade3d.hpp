// ge_facade3d.hpp -- C++ facade over the C ABI of ge_physics3d.h that keeps
// the reference's 3D call names, so mode3D() (GE/src/globalLogic.cpp:
// 1552-1722) can call the device path stage by stage:
//
//   reference                                         facade
//   glm::vec4 boundingBox3D()        globalLogic.cpp:1465   ge_gpu::boundingBox3D(scene)
//   Morton::computeMortonKeys3D      morton.cpp:94-113      ge_gpu::Morton3D::computeMortonKeys3D
//   Morton::sortParticlesByMortonKey3D  morton.cpp:115-137  ge_gpu::Morton3D::sortParticlesByMortonKey3D
//   Octree root3D(pP, rP, pos, size) quadtree.cpp:275-292   ge_gpu::Octree(scene, bb)
//   Physics3D::calculateForceFromGrid3D physics3D.cpp:131   ge_gpu::Physics3D::calculateForceFromGrid3D
//   Physics3D::naiveGravity3D        physics3D.cpp:416      ge_gpu::Physics3D::naiveGravity3D
//   Physics3D::integrateStart3D / End3D physics3D.cpp:740, 1158  ge_gpu::Physics3D::integrateStart3D / End3D
//   NeighborSearch3DV2AVX2::newGridAVX2 / neighborAmount  neighborSearch.h:1017, 1175
//                                                         ge_gpu::NeighborSearch3DV2AVX2
//   SPH3D::pcisphSolver              SPH3D.h:590            ge_gpu::SPH3D::pcisphSolver
//
// As in ge_facade.hpp, the particles live in a Scene3D (one ge3_ctx) and the
// physics variables are a ge3_step_params; failures throw ge_gpu::Error.
// Header-only; link libge_physics.so.
#ifndef GE_FACADE3D_HPP
#define GE_FACADE3D_HPP

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ge_physics3d.h"

namespace ge_gpu {

#ifndef GE_GPU_ERROR_DEFINED
#define GE_GPU_ERROR_DEFINED
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
#endif

// Host-side structure of arrays of the ParticlePhysics3D fields on the path
// (GE/include/Particles/particle.h:142-212); empty vectors mean "not given".
struct Particles3D {
    std::vector<float> x, y, z, vx, vy, vz, ax, ay, az, pvx, pvy, pvz, mass;
    std::vector<uint32_t> flags, id;
    std::vector<float> sph_mass, rest_dens, stiff, visc, cohesion, dens, press, press_tmp, pred_dens;
    std::vector<float> ppx, ppy, ppz, pfx, pfy, pfz;

    size_t size() const { return x.size(); }
    std::vector<std::vector<float> *> floats() {
        return {&x, &y, &z, &vx, &vy, &vz, &ax, &ay, &az, &pvx, &pvy, &pvz, &mass, &sph_mass, &rest_dens, &stiff,
                &visc, &cohesion, &dens, &press, &press_tmp, &pred_dens, &ppx, &ppy, &ppz, &pfx, &pfy, &pfz};
    }
    ge3_particles view() {
        auto f = [](std::vector<float> &v) { return v.empty() ? nullptr : v.data(); };
        auto u = [](std::vector<uint32_t> &v) { return v.empty() ? nullptr : v.data(); };
        ge3_particles p{};
        p.x = f(x); p.y = f(y); p.z = f(z); p.vx = f(vx); p.vy = f(vy); p.vz = f(vz);
        p.ax = f(ax); p.ay = f(ay); p.az = f(az); p.pvx = f(pvx); p.pvy = f(pvy); p.pvz = f(pvz);
        p.mass = f(mass); p.flags = u(flags); p.id = u(id);
        p.sph_mass = f(sph_mass); p.rest_dens = f(rest_dens); p.stiff = f(stiff); p.visc = f(visc);
        p.cohesion = f(cohesion); p.dens = f(dens); p.press = f(press); p.press_tmp = f(press_tmp);
        p.pred_dens = f(pred_dens); p.ppx = f(ppx); p.ppy = f(ppy); p.ppz = f(ppz);
        p.pfx = f(pfx); p.pfy = f(pfy); p.pfz = f(pfz);
        return p;
    }
};

// One 3D scene resident on one device (the reference's myParam.pParticles3D).
class Scene3D {
public:
    explicit Scene3D(uint32_t capacity, bool sph = false, int device = 0) {
        ge3_config cfg;
        ge3_default_config(&cfg);
        cfg.capacity = capacity;
        cfg.device = device;
        cfg.enable_sph = sph ? 1 : 0;
        const int r = ge3_create(&ctx_, &cfg);
        if (r != GE_OK) throw Error(r, "ge3_create failed (status " + std::to_string(r) + ")");
    }
    ~Scene3D() {
        if (ctx_) ge3_destroy(ctx_);
    }
    Scene3D(const Scene3D &) = delete;
    Scene3D &operator=(const Scene3D &) = delete;

    ge3_ctx *ctx() const { return ctx_; }
    void check(int r) const {
        if (r != GE_OK) throw Error(r, ge3_last_error(ctx_));
    }
    void upload(Particles3D &p) {
        ge3_particles v = p.view();
        check(ge3_upload(ctx_, (uint32_t)p.size(), &v));
    }
    // fills every non-empty vector of p (resized to the particle count) in the device order
    void download(Particles3D &p) {
        uint32_t n = 0;
        check(ge3_count(ctx_, &n));
        for (std::vector<float> *v : p.floats())
            if (!v->empty()) v->resize(n);
        if (!p.flags.empty()) p.flags.resize(n);
        if (!p.id.empty()) p.id.resize(n);
        ge3_particles v = p.view();
        check(ge3_download(ctx_, &v));
    }

private:
    ge3_ctx *ctx_ = nullptr;
};

// glm::vec4 boundingBox3D() -- {x, y, z, size} of the cube around all particles
inline std::array<float, 4> boundingBox3D(Scene3D &s) {
    std::array<float, 4> bb{};
    s.check(ge3_bounding_box(s.ctx(), bb.data()));
    return bb;
}

struct Morton3D {
    void computeMortonKeys3D(Scene3D &s, const std::array<float, 4> &bb) {
        s.check(ge3_compute_morton_keys(s.ctx(), bb.data()));
    }
    void sortParticlesByMortonKey3D(Scene3D &s) { s.check(ge3_sort_by_morton_key(s.ctx())); }
};

// Octree root3D(pParticles3D, rParticles3D, pos, size) over the sorted particles
struct Octree {
    Octree(Scene3D &s, const std::array<float, 4> &bb) : scene(&s) { s.check(ge3_build_octree(s.ctx(), bb.data())); }
    std::vector<ge3_node> nodes() const {
        uint32_t nn = 0;
        scene->check(ge3_get_node_count(scene->ctx(), &nn));
        std::vector<ge3_node> out(nn);
        scene->check(ge3_get_nodes(scene->ctx(), out.data(), nn));
        return out;
    }
    Scene3D *scene;
};

struct Physics3D {
    // acc = 0, then the octree walk (or the direct sum by mode3D's dispatch)
    void calculateForceFromGrid3D(Scene3D &s, const ge3_step_params &myVar) {
        s.check(ge3_compute_gravity(s.ctx(), &myVar));
    }
    void naiveGravity3D(Scene3D &s, const ge3_step_params &myVar) {
        ge3_step_params p = myVar;
        p.naive = 1;
        s.check(ge3_compute_gravity(s.ctx(), &p));
    }
    void integrateStart3D(Scene3D &s, const ge3_step_params &myVar) { s.check(ge3_integrate_start(s.ctx(), &myVar)); }
    void integrateEnd3D(Scene3D &s, const ge3_step_params &myVar) { s.check(ge3_integrate_end(s.ctx(), &myVar)); }
};

struct NeighborSearch3DV2AVX2 {
    void newGridAVX2(Scene3D &s, const ge3_step_params &myVar) { s.check(ge3_sph_new_grid(s.ctx(), &myVar)); }
    // rParticles3D[i].neighbors for every particle, in the device order
    std::vector<uint32_t> neighborAmount(Scene3D &s, const ge3_step_params &myVar) {
        uint32_t n = 0;
        s.check(ge3_count(s.ctx(), &n));
        std::vector<uint32_t> out(n);
        s.check(ge3_neighbor_counts(s.ctx(), &myVar, out.data()));
        return out;
    }
};

struct SPH3D {
    void pcisphSolver(Scene3D &s, const ge3_step_params &myVar) { s.check(ge3_sph_pcisph(s.ctx(), &myVar)); }
};

// The physics part of mode3D() for `frames` frames, fused on the device.
inline void mode3D(Scene3D &s, const ge3_step_params &myVar, int frames = 1) {
    s.check(ge3_step(s.ctx(), &myVar, frames));
    s.check(ge3_sync(s.ctx()));
}

}  // namespace ge_gpu

#endif  // GE_FACADE3D_HPP
