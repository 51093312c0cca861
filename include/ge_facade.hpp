// ge_facade.hpp -- C++ facade over the C ABI of ge_physics.h that keeps the
// reference's call names, so the 2D step of updateScene()
// (GE/src/globalLogic.cpp:1155-1372) can call the device path stage by stage:
//
//   reference (GE/include/Physics/...)                   facade
//   glm::vec3 boundingBox()             globalLogic.cpp:1083   ge_gpu::boundingBox(scene)
//   Morton::computeMortonKeys           morton.h:17            ge_gpu::Morton::computeMortonKeys
//   Morton::sortParticlesByMortonKey    morton.h:19-20         ge_gpu::Morton::sortParticlesByMortonKey
//   Quadtree::Quadtree(pP, rP, bb)      quadtree.h:118-125     ge_gpu::Quadtree(scene, bb)
//   Physics::calculateForceFromGrid     physics.h:29           ge_gpu::Physics::calculateForceFromGrid
//   Physics::naiveGravity               physics.h:48           ge_gpu::Physics::naiveGravity
//   Physics::integrateStart / End       physics.h:66-68        ge_gpu::Physics::integrateStart / End
//   NeighborSearchV2AVX2::newGridAVX2   neighborSearch.h:636   ge_gpu::NeighborSearchV2AVX2::newGridAVX2
//   SPH::pcisphSolver                   SPH.h:623              ge_gpu::SPH::pcisphSolver
//
// The reference's functions work in place on std::vector<ParticlePhysics>
// and global state; here the particles live in a Scene (one ge_ctx: device
// memory and a stream) and the physics variables are a ge_step_params.  The
// reference reports no errors; the facade throws ge_gpu::Error with the
// context's message when a call fails.  Header-only; link libge_physics.so.
#ifndef GE_FACADE_HPP
#define GE_FACADE_HPP

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ge_physics.h"

namespace ge_gpu {

#ifndef GE_GPU_ERROR_DEFINED
#define GE_GPU_ERROR_DEFINED
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
#endif

// Host-side structure of arrays of the ParticlePhysics fields on the path
// (GE/include/Particles/particle.h:5-73); empty vectors mean "not given".
struct Particles {
    std::vector<float> x, y, vx, vy, ax, ay, pvx, pvy, mass;
    std::vector<uint32_t> flags, id;
    std::vector<float> sph_mass, rest_dens, stiff, visc, cohesion, dens;
    std::vector<float> press, press_tmp, pred_dens, ppx, ppy, pfx, pfy;

    size_t size() const { return x.size(); }
    // a ge_particles view (empty vectors become NULL)
    ge_particles view() {
        auto f = [](std::vector<float> &v) { return v.empty() ? nullptr : v.data(); };
        auto u = [](std::vector<uint32_t> &v) { return v.empty() ? nullptr : v.data(); };
        ge_particles p{};
        p.x = f(x); p.y = f(y); p.vx = f(vx); p.vy = f(vy); p.ax = f(ax); p.ay = f(ay);
        p.pvx = f(pvx); p.pvy = f(pvy); p.mass = f(mass); p.flags = u(flags); p.id = u(id);
        p.sph_mass = f(sph_mass); p.rest_dens = f(rest_dens); p.stiff = f(stiff); p.visc = f(visc);
        p.cohesion = f(cohesion); p.dens = f(dens); p.press = f(press); p.press_tmp = f(press_tmp);
        p.pred_dens = f(pred_dens); p.ppx = f(ppx); p.ppy = f(ppy); p.pfx = f(pfx); p.pfy = f(pfy);
        return p;
    }
};

// One scene: the particle state resident on one device (the reference's
// myParam.pParticles, globalLogic.cpp:3, moved to HBM).
class Scene {
public:
    explicit Scene(uint32_t capacity, bool sph = false, int device = 0) {
        ge_config cfg;
        ge_default_config(&cfg);
        cfg.capacity = capacity;
        cfg.device = device;
        cfg.enable_sph = sph ? 1 : 0;
        const int r = ge_create(&ctx_, &cfg);
        if (r != GE_OK) throw Error(r, "ge_create failed (status " + std::to_string(r) + ")");
        sph_ = sph;
    }
    ~Scene() {
        if (ctx_) ge_destroy(ctx_);
    }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;

    ge_ctx *ctx() const { return ctx_; }
    bool sph() const { return sph_; }
    void check(int r) const {
        if (r != GE_OK) throw Error(r, ge_last_error(ctx_));
    }
    void upload(Particles &p) {
        ge_particles v = p.view();
        check(ge_upload(ctx_, (uint32_t)p.size(), &v));
    }
    // fills every non-empty vector of p (resized to the particle count), in
    // the device order (Morton order after a sort, as the reference leaves it)
    void download(Particles &p) {
        uint32_t n = 0;
        check(ge_count(ctx_, &n));
        auto fit = [n](auto &v) {
            if (!v.empty()) v.resize(n);
        };
        fit(p.x); fit(p.y); fit(p.vx); fit(p.vy); fit(p.ax); fit(p.ay); fit(p.pvx); fit(p.pvy); fit(p.mass);
        fit(p.flags); fit(p.id); fit(p.sph_mass); fit(p.rest_dens); fit(p.stiff); fit(p.visc);
        fit(p.cohesion); fit(p.dens); fit(p.press); fit(p.press_tmp); fit(p.pred_dens); fit(p.ppx);
        fit(p.ppy); fit(p.pfx); fit(p.pfy);
        ge_particles v = p.view();
        check(ge_download(ctx_, &v));
    }
    uint32_t size() const {
        uint32_t n = 0;
        check(ge_count(ctx_, &n));
        return n;
    }

private:
    ge_ctx *ctx_ = nullptr;
    bool sph_ = false;
};

// glm::vec3 boundingBox() -- {x, y, size} of the square around all particles
inline std::array<float, 3> boundingBox(Scene &s) {
    std::array<float, 3> bb{};
    s.check(ge_bounding_box(s.ctx(), bb.data()));
    return bb;
}

struct Morton {
    void computeMortonKeys(Scene &s, const std::array<float, 3> &posSize) {
        s.check(ge_compute_morton_keys(s.ctx(), posSize.data()));
    }
    void sortParticlesByMortonKey(Scene &s) { s.check(ge_sort_by_morton_key(s.ctx())); }
};

// Quadtree root(pParticles, rParticles, bb): builds the tree of the scene's
// current (sorted) particles; nodes() reads it back in the reference layout.
struct Quadtree {
    Quadtree(Scene &s, const std::array<float, 3> &bb) : scene(&s) {
        s.check(ge_build_quadtree(s.ctx(), bb.data()));
    }
    std::vector<ge_node> nodes() const {
        uint32_t nn = 0;
        scene->check(ge_get_node_count(scene->ctx(), &nn));
        std::vector<ge_node> out(nn);
        scene->check(ge_get_nodes(scene->ctx(), out.data(), nn));
        return out;
    }
    Scene *scene;
};

struct Physics {
    // acc = 0, then the tree walk (or the direct sum by the reference's
    // dispatch, globalLogic.cpp:1319-1348)
    void calculateForceFromGrid(Scene &s, const ge_step_params &myVar) {
        s.check(ge_compute_gravity(s.ctx(), &myVar));
    }
    void naiveGravity(Scene &s, const ge_step_params &myVar) {
        ge_step_params p = myVar;
        p.naive = 1;
        s.check(ge_compute_gravity(s.ctx(), &p));
    }
    void integrateStart(Scene &s, const ge_step_params &myVar) { s.check(ge_integrate_start(s.ctx(), &myVar)); }
    void integrateEnd(Scene &s, const ge_step_params &myVar) { s.check(ge_integrate_end(s.ctx(), &myVar)); }
};

struct NeighborSearchV2AVX2 {
    void newGridAVX2(Scene &s, const ge_step_params &myVar) { s.check(ge_sph_new_grid(s.ctx(), &myVar)); }
};

struct SPH {
    void pcisphSolver(Scene &s, const ge_step_params &myVar) { s.check(ge_sph_pcisph(s.ctx(), &myVar)); }
};

// The whole 2D physics part of updateScene() for `frames` frames, fused on
// the device (one launch sequence per frame, no host round trip).
inline void updateScene(Scene &s, const ge_step_params &myVar, int frames = 1) {
    s.check(ge_step(s.ctx(), &myVar, frames));
    s.check(ge_sync(s.ctx()));
}

}  // namespace ge_gpu

#endif  // GE_FACADE_HPP
