// ge_host.cpp -- the autotuner and the continuous-batching request server
// (include/ge_serve.h).  Host C++ only: every device operation goes through
// the C ABI of ge_physics.h, so this file is the kind of caller the ABI is
// written for.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ge_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Host copy of every field a context may hold (SPH fields are ignored by
// contexts without SPH state).
struct HostState {
    std::vector<float> f[22];
    std::vector<uint32_t> flags, id;
    ge_particles p{};

    explicit HostState(uint32_t n) {
        for (auto &v : f) v.resize(n ? n : 1);
        flags.resize(n ? n : 1);
        id.resize(n ? n : 1);
        float **fp[22] = {&p.x, &p.y, &p.vx, &p.vy, &p.ax, &p.ay, &p.pvx, &p.pvy, &p.mass,
                          &p.sph_mass, &p.rest_dens, &p.stiff, &p.visc, &p.cohesion, &p.dens, &p.press,
                          &p.press_tmp, &p.pred_dens, &p.ppx, &p.ppy, &p.pfx, &p.pfy};
        for (int k = 0; k < 22; ++k) *fp[k] = f[k].data();
        p.flags = flags.data();
        p.id = id.data();
    }
};

int bits_for(uint64_t v) {  // bits to hold values 0 .. v-1 (as the batched sort word does)
    int b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

}  // namespace

// ------------------------------------------------- array-of-structures --

namespace {

// Field k of ge_particles (declaration order) as a pointer slot.
void **field_slot(ge_particles &p, int k) {
    void **slots[GE_NFIELDS] = {
        (void **)&p.x, (void **)&p.y, (void **)&p.vx, (void **)&p.vy, (void **)&p.ax, (void **)&p.ay,
        (void **)&p.pvx, (void **)&p.pvy, (void **)&p.mass, (void **)&p.flags, (void **)&p.id,
        (void **)&p.sph_mass, (void **)&p.rest_dens, (void **)&p.stiff, (void **)&p.visc, (void **)&p.cohesion,
        (void **)&p.dens, (void **)&p.press, (void **)&p.press_tmp, (void **)&p.pred_dens, (void **)&p.ppx,
        (void **)&p.ppy, (void **)&p.pfx, (void **)&p.pfy};
    return slots[k];
}

constexpr int kFieldFlags = 9;

bool layout_ok(const ge_aos_layout *l) {
    if (!l || l->stride == 0) return false;
    for (int k = 0; k < GE_NFIELDS; ++k)
        if (l->off[k] >= 0 && (uint32_t)l->off[k] + 4u > l->stride) return false;
    return true;
}

}  // namespace

extern "C" int ge_upload_aos(ge_ctx *ctx, uint32_t n, const void *records, const ge_aos_layout *layout,
                             const uint32_t *flags) {
    if (!ctx || (n && !records) || !layout_ok(layout)) return GE_E_INVALID;
    // flattenParticles: one SoA column per present field
    std::vector<std::vector<uint32_t>> cols(GE_NFIELDS);
    ge_particles p{};
    const unsigned char *base = static_cast<const unsigned char *>(records);
    for (int k = 0; k < GE_NFIELDS; ++k) {
        if (k == kFieldFlags || layout->off[k] < 0) continue;
        cols[k].resize(n ? n : 1);
        for (uint32_t i = 0; i < n; ++i)
            std::memcpy(&cols[k][i], base + (size_t)i * layout->stride + layout->off[k], 4);
        *field_slot(p, k) = cols[k].data();
    }
    p.flags = const_cast<uint32_t *>(flags);
    return ge_upload(ctx, n, &p);
}

extern "C" int ge_download_aos(ge_ctx *ctx, void *records, const ge_aos_layout *layout, uint32_t *flags) {
    uint32_t n = 0;
    if (!ctx || !layout_ok(layout)) return GE_E_INVALID;
    if (int r = ge_count(ctx, &n)) return r;
    if (n && !records) return GE_E_INVALID;
    std::vector<std::vector<uint32_t>> cols(GE_NFIELDS);
    ge_particles p{};
    for (int k = 0; k < GE_NFIELDS; ++k) {
        if (k == kFieldFlags || layout->off[k] < 0) continue;
        cols[k].resize(n ? n : 1);
        *field_slot(p, k) = cols[k].data();
    }
    p.flags = flags;
    if (int r = ge_download(ctx, &p)) return r;
    // readFlattenBack: every present field back into its record
    unsigned char *base = static_cast<unsigned char *>(records);
    for (int k = 0; k < GE_NFIELDS; ++k) {
        if (k == kFieldFlags || layout->off[k] < 0) continue;
        for (uint32_t i = 0; i < n; ++i)
            std::memcpy(base + (size_t)i * layout->stride + layout->off[k], &cols[k][i], 4);
    }
    return GE_OK;
}

// ------------------------------------------------------------------ tuner --

extern "C" int ge_autotune(ge_ctx *ctx, const ge_step_params *base, int32_t frames, ge_tune_result *out) {
    if (!ctx || !base || !out || frames < 1) return GE_E_INVALID;
    uint32_t n = 0, nsys = 0;
    if (int r = ge_count(ctx, &n)) return r;
    if (n == 0) return GE_E_INVALID;
    if (int r = ge_system_offsets(ctx, &nsys, nullptr, 0)) return r;
    std::vector<uint32_t> off(nsys + 1), counts(nsys);
    if (int r = ge_system_offsets(ctx, &nsys, off.data(), nsys + 1)) return r;
    for (uint32_t s = 0; s < nsys; ++s) counts[s] = off[s + 1] - off[s];
    HostState snap(n);
    if (int r = ge_download(ctx, &snap.p)) return r;
    auto restore = [&]() { return ge_upload_systems(ctx, nsys, counts.data(), &snap.p); };
    auto time_one = [&](uint32_t leaf, uint32_t block, float *ms) -> int {
        ge_step_params p = *base;
        p.leaf_capacity = leaf;
        p.walk_block = block;
        p.dt_start = p.dt;
        if (int r = restore()) return r;
        if (int r = ge_step(ctx, &p, 1)) return r;
        if (int r = ge_sync(ctx)) return r;
        const Clock::time_point t0 = Clock::now();
        if (int r = ge_step(ctx, &p, frames)) return r;
        if (int r = ge_sync(ctx)) return r;
        *ms = (float)(seconds_since(t0) * 1e3 / frames);
        return GE_OK;
    };
    static const uint32_t kLeaf[3] = {8, 16, 32}, kBlockSz[3] = {64, 128, 256};
    ge_tune_result res{};
    float best = 0.0f;
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) {
            float ms = 0.0f;
            if (int r = time_one(kLeaf[a], kBlockSz[b], &ms)) {
                (void)restore();
                return r;
            }
            res.table_ms[a][b] = ms;
            if (best == 0.0f || ms < best) {
                best = ms;
                res.leaf_capacity = kLeaf[a];
                res.walk_block = kBlockSz[b];
                res.ms_per_frame = ms;
            }
        }
    }
    if (int r = time_one(0, 256, &res.reference_ms)) {
        (void)restore();
        return r;
    }
    if (int r = restore()) return r;
    *out = res;
    return GE_OK;
}

// ----------------------------------------------------------------- server --

struct ge_server {
    ge_server_config cfg{};
    ge_step_params pg{}, ps{};
    ge_ctx *batch = nullptr;
    uint32_t batch_cap = 0;
    struct SphSlot {
        ge_ctx *ctx = nullptr;
        int64_t req = -1;
        int32_t remaining = 0;
        int32_t inflight = 0;
    };
    std::vector<SphSlot> sph;
    uint64_t device_bytes = 0;
    uint64_t round_limit = 0;  // ge_server_set_round_limit (0: none)
    std::string err;
};

namespace {

int server_fail(ge_server *s, int code, const std::string &msg) {
    if (s) s->err = msg;
    return code;
}

// Returns every context to the idle state a run starts from: waits for the
// chunks and rounds still queued, forgets which request each SPH context held,
// and empties the batch.  Runs at the start of every ge_server_run and on each
// of its error returns, so a run that ended early leaves no slot that a later
// run would take for busy (or finish with a request index of the old run).
void server_reset(ge_server *s) {
    for (auto &slot : s->sph) {
        if (slot.ctx) (void)ge_sync(slot.ctx);
        slot.req = -1;
        slot.remaining = 0;
        slot.inflight = 0;
    }
    if (s->batch) {
        (void)ge_sync(s->batch);
        (void)ge_systems_update(s->batch, 0, nullptr, 0, nullptr, nullptr);
    }
}

int server_abort(ge_server *s, int code, const std::string &msg) {
    server_reset(s);
    return server_fail(s, code, msg);
}

// Particles of a gravity request staged for one ge_systems_update call.
struct Staging {
    std::vector<float> x, y, vx, vy, ax, ay, pvx, pvy, m;
    std::vector<uint32_t> flags, id;
    void add(const ge_request &r) {
        auto put = [&](std::vector<float> &dst, const float *src) {
            if (src) dst.insert(dst.end(), src, src + r.n);
            else dst.insert(dst.end(), r.n, 0.0f);
        };
        put(x, r.in.x); put(y, r.in.y); put(vx, r.in.vx); put(vy, r.in.vy);
        put(ax, r.in.ax); put(ay, r.in.ay); put(pvx, r.in.pvx); put(pvy, r.in.pvy); put(m, r.in.mass);
        if (r.in.flags) flags.insert(flags.end(), r.in.flags, r.in.flags + r.n);
        else flags.insert(flags.end(), r.n, 0u);
        if (r.in.id) {
            id.insert(id.end(), r.in.id, r.in.id + r.n);
        } else {
            for (uint32_t i = 0; i < r.n; ++i) id.push_back(i);
        }
    }
    ge_particles view() {
        ge_particles p{};
        p.x = x.data(); p.y = y.data(); p.vx = vx.data(); p.vy = vy.data();
        p.ax = ax.data(); p.ay = ay.data(); p.pvx = pvx.data(); p.pvy = pvy.data(); p.mass = m.data();
        p.flags = flags.data(); p.id = id.data();
        return p;
    }
};

}  // namespace

extern "C" {

void ge_server_default_config(ge_server_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->device = 0;
    cfg->hbm_budget_bytes = 200ull << 30;  // 200 GB of the 288 GB per MI355X
    cfg->batch_capacity = 1u << 26;
    cfg->max_systems = 512;
    cfg->sph_slots = 8;
    cfg->sph_capacity = 1u << 19;
    cfg->round_frames = 8;
    cfg->batching = 1;
    cfg->use_graph = 1;
}

int ge_server_create(ge_server **out, const ge_server_config *cfg, const ge_step_params *gravity,
                     const ge_step_params *sph) {
    if (!out || !cfg || !gravity) return GE_E_INVALID;
    *out = nullptr;
    if (cfg->sph_slots && !sph) return GE_E_INVALID;
    if (cfg->max_systems == 0 || cfg->max_systems > 2048 || cfg->round_frames == 0) return GE_E_INVALID;
    ge_server *s = new (std::nothrow) ge_server();
    if (!s) return GE_E_NOMEM;
    s->cfg = *cfg;
    // every request runs exactly as it would alone: fixed timestep, tree walk
    // for every size (the batched form never takes the direct-sum branch)
    s->pg = *gravity;
    s->pg.dt_start = s->pg.dt;
    s->pg.naive = 0;
    s->pg.direct_below = 0;
    s->pg.sph = 0;
    s->pg.use_graph = cfg->use_graph;
    if (sph) {
        s->ps = *sph;
        s->ps.dt_start = s->ps.dt;
        s->ps.naive = 0;
        s->ps.direct_below = 0;
        s->ps.sph = 1;
        s->ps.use_graph = cfg->use_graph;
    }
    uint64_t budget = cfg->hbm_budget_bytes;
    for (uint32_t k = 0; k < cfg->sph_slots; ++k) {
        ge_config c;
        ge_default_config(&c);
        c.capacity = cfg->sph_capacity;
        c.device = cfg->device;
        c.enable_sph = 1;
        ge_server::SphSlot slot;
        if (int r = ge_create(&slot.ctx, &c)) {
            ge_server_destroy(s);
            return r;
        }
        uint64_t b = 0;
        ge_memory_bytes(slot.ctx, &b);
        s->sph.push_back(slot);
        s->device_bytes += b;
        if (s->device_bytes > budget) {
            ge_server_destroy(s);
            return GE_E_NOMEM;
        }
    }
    // batch capacity: the configured bound, shrunk until the context fits
    // what the SPH contexts left of the budget
    uint64_t cap = cfg->batch_capacity;
    {
        ge_config c;
        ge_default_config(&c);
        c.capacity = 1u << 16;
        c.device = cfg->device;
        c.max_systems = cfg->max_systems;
        ge_ctx *probe = nullptr;
        if (int r = ge_create(&probe, &c)) {
            ge_server_destroy(s);
            return r;
        }
        uint64_t pb = 0;
        ge_memory_bytes(probe, &pb);
        ge_destroy(probe);
        const uint64_t per = pb / c.capacity + 1;
        const uint64_t left = budget - s->device_bytes;
        if (cap > left / per) cap = left / per;
    }
    while (cap >= 1024) {
        ge_config c;
        ge_default_config(&c);
        c.capacity = (uint32_t)cap;
        c.device = cfg->device;
        c.max_systems = cfg->max_systems;
        if (ge_create(&s->batch, &c) == GE_OK) {
            uint64_t b = 0;
            ge_memory_bytes(s->batch, &b);
            if (s->device_bytes + b <= budget) {
                s->device_bytes += b;
                break;
            }
            ge_destroy(s->batch);
            s->batch = nullptr;
        }
        cap = cap * 9 / 10;
    }
    if (!s->batch) {
        ge_server_destroy(s);
        return GE_E_NOMEM;
    }
    s->batch_cap = (uint32_t)cap;
    *out = s;
    return GE_OK;
}

int ge_server_destroy(ge_server *s) {
    if (!s) return GE_E_INVALID;
    for (auto &slot : s->sph)
        if (slot.ctx) ge_destroy(slot.ctx);
    if (s->batch) ge_destroy(s->batch);
    delete s;
    return GE_OK;
}

const char *ge_server_last_error(const ge_server *s) { return s ? s->err.c_str() : "null server"; }

int ge_server_set_round_limit(ge_server *s, uint64_t max_rounds) {
    if (!s) return GE_E_INVALID;
    s->round_limit = max_rounds;
    return GE_OK;
}

int ge_server_run(ge_server *s, uint32_t nreq, const ge_request *reqs, ge_request_result *res,
                  ge_server_stats *stats) {
    if (!s || (nreq && (!reqs || !res))) return GE_E_INVALID;
    const ge_server_config &cfg = s->cfg;
    std::vector<uint32_t> order(nreq);
    for (uint32_t i = 0; i < nreq; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    for (uint32_t i = 0; i < nreq; ++i) {
        res[i].arrival_s = reqs[i].arrival_s;
        res[i].admit_s = res[i].done_s = -1.0;
        res[i].status = GE_OK;
    }
    struct Resident {
        uint32_t req;
        int32_t remaining;
    };
    std::vector<Resident> resident;  // batch systems, in batch order
    std::deque<uint32_t> waiting;
    uint32_t next = 0, done = 0, max_resident = 0;
    uint64_t rounds = 0;
    // start from idle contexts and an empty batch, whatever the last run left
    server_reset(s);
    if (s->batch) {
        uint32_t left = 1;
        if (ge_count(s->batch, &left) != GE_OK || left != 0)
            return server_abort(s, GE_E_HIP, "the batch context could not be emptied");
    }
    const Clock::time_point t0 = Clock::now();
    auto finish = [&](uint32_t q, int status) {
        res[q].done_s = seconds_since(t0);
        res[q].status = status;
        ++done;
    };
    auto sph_busy = [&]() {
        for (auto &slot : s->sph)
            if (slot.req >= 0) return true;
        return false;
    };
    // The batch round and the SPH chunks run asynchronously; the loop wakes
    // on whichever finishes first, so an SPH slot is refilled as soon as its
    // chunk ends rather than when the (longer) batch round does.  The batch
    // changes membership only between rounds.
    int32_t batch_R = 0;  // frames of the batch round in flight (0: the batch is idle)
    while (done < nreq) {
        if (s->round_limit && rounds >= s->round_limit)
            return server_abort(s, GE_E_CAPACITY, "round limit reached");
        const double now = seconds_since(t0);
        while (next < nreq && reqs[order[next]].arrival_s <= now) waiting.push_back(order[next++]);
        const bool batch_idle = batch_R == 0;

        // ---- admission: first come, first served per kind, within capacity
        std::vector<uint32_t> keep;
        std::vector<Resident> next_res;
        if (batch_idle) {
            for (uint32_t t = 0; t < resident.size(); ++t)
                if (resident[t].remaining > 0) keep.push_back(t);
            for (uint32_t t : keep) next_res.push_back(resident[t]);
        } else {
            next_res = resident;
        }
        uint64_t used = 0;
        uint32_t maxn = 0;
        for (auto &r : next_res) {
            used += reqs[r.req].n;
            maxn = std::max(maxn, reqs[r.req].n);
        }
        Staging stage;
        std::vector<uint32_t> add_counts;
        const bool idle = batch_idle && next_res.empty() && !sph_busy();
        bool grav_blocked = !batch_idle, sph_blocked = false, took = false;
        for (auto it = waiting.begin(); it != waiting.end();) {
            const uint32_t q = *it;
            const ge_request &r = reqs[q];
            if (r.n == 0 || r.steps <= 0 || (r.sph && s->sph.empty()) ||
                (!r.sph && r.n > s->batch_cap) || (r.sph && r.n > cfg.sph_capacity) || !r.in.x || !r.in.y ||
                !r.in.mass) {
                res[q].admit_s = seconds_since(t0);
                finish(q, r.n == 0 || r.steps <= 0 || !r.in.x ? GE_E_INVALID : GE_E_CAPACITY);
                it = waiting.erase(it);
                continue;
            }
            if (!cfg.batching) {
                // plain form: one request on the device at a time, in order
                if (!idle || took) break;
            }
            bool admitted = false;
            if (r.sph && !sph_blocked) {
                for (auto &slot : s->sph) {
                    if (slot.req >= 0) continue;
                    const int rc = ge_upload(slot.ctx, r.n, &r.in);
                    res[q].admit_s = seconds_since(t0);
                    if (rc != GE_OK) {
                        finish(q, rc);
                    } else {
                        slot.req = q;
                        slot.remaining = r.steps;
                        slot.inflight = 0;
                    }
                    admitted = true;
                    break;
                }
                if (!admitted) sph_blocked = true;
            } else if (!r.sph && !grav_blocked) {
                const uint32_t nsys = (uint32_t)next_res.size() + 1;
                const uint32_t mx = std::max(maxn, r.n);
                const bool fits = used + r.n <= s->batch_cap && nsys <= cfg.max_systems &&
                                  (nsys == 1 || bits_for(mx) + bits_for(nsys) <= 28);
                if (fits) {
                    stage.add(r);
                    add_counts.push_back(r.n);
                    next_res.push_back(Resident{q, r.steps});
                    used += r.n;
                    maxn = mx;
                    res[q].admit_s = seconds_since(t0);
                    admitted = true;
                } else {
                    grav_blocked = true;
                }
            }
            if (admitted) {
                took = true;
                it = waiting.erase(it);
            } else {
                if (!cfg.batching) break;
                ++it;
            }
            if (grav_blocked && sph_blocked) break;
        }
        if (batch_idle && (keep.size() != resident.size() || !add_counts.empty())) {
            ge_particles pv = stage.view();
            const int rc = ge_systems_update(s->batch, (uint32_t)keep.size(), keep.data(),
                                             (uint32_t)add_counts.size(), add_counts.data(), &pv);
            if (rc != GE_OK) return server_abort(s, rc, ge_last_error(s->batch));
            resident.swap(next_res);
        }
        uint32_t on_device = (uint32_t)resident.size();
        for (auto &slot : s->sph) on_device += slot.req >= 0;
        max_resident = std::max(max_resident, on_device);

        // ---- launch: a batch round if the batch is idle, the next chunk of every idle SPH slot
        if (batch_idle && !resident.empty()) {
            int32_t R = (int32_t)cfg.round_frames;
            for (auto &r : resident) R = std::min(R, r.remaining);
            if (int rc = ge_step(s->batch, &s->pg, R)) return server_abort(s, rc, ge_last_error(s->batch));
            batch_R = R;
            ++rounds;
        }
        for (auto &slot : s->sph) {
            if (slot.req < 0 || slot.inflight) continue;
            const int32_t rs = std::min((int32_t)cfg.round_frames, slot.remaining);
            const int rc = ge_step(slot.ctx, &s->ps, rs);
            if (rc != GE_OK) {
                finish((uint32_t)slot.req, rc);
                slot.req = -1;
                continue;
            }
            slot.inflight = rs;
        }

        // ---- wait for the first completion (or an arrival while work runs)
        uint32_t inflight = batch_R > 0 ? 1u : 0u;
        ge_server::SphSlot *only = nullptr;
        for (auto &slot : s->sph)
            if (slot.req >= 0 && slot.inflight) {
                ++inflight;
                only = &slot;
            }
        if (inflight == 1 && (cfg.batching == 0 || waiting.empty())) {
            // one context on the device and nothing that could start beside
            // it: block on it
            if (batch_R > 0) {
                if (int rc = ge_sync(s->batch)) return server_abort(s, rc, ge_last_error(s->batch));
            } else {
                (void)ge_sync(only->ctx);  // its status is read with the completion below
            }
        } else if (inflight > 0) {
            bool any = false;
            while (!any) {
                int32_t d = 0;
                if (batch_R > 0 && ge_poll(s->batch, &d) == GE_OK && d) any = true;
                for (auto &slot : s->sph) {
                    d = 0;
                    if (slot.req >= 0 && slot.inflight && ge_poll(slot.ctx, &d) == GE_OK && d) any = true;
                }
                if (!any) {
                    if (next < nreq && reqs[order[next]].arrival_s <= seconds_since(t0)) break;
                    std::this_thread::sleep_for(std::chrono::microseconds(50));
                }
            }
        } else if (waiting.empty() && next < nreq) {
            const double wait = reqs[order[next]].arrival_s - seconds_since(t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        }

        // ---- completions
        if (batch_R > 0) {
            int32_t d = 0;
            if (ge_poll(s->batch, &d) == GE_OK && d) {
                if (int rc = ge_sync(s->batch)) return server_abort(s, rc, ge_last_error(s->batch));
                for (auto &r : resident) r.remaining -= batch_R;
                batch_R = 0;
                for (uint32_t t = 0; t < resident.size(); ++t) {
                    if (resident[t].remaining > 0) continue;
                    const uint32_t q = resident[t].req;
                    finish(q, ge_download_system(s->batch, t, &reqs[q].out));
                }
            }
        }
        for (auto &slot : s->sph) {
            if (slot.req < 0 || !slot.inflight) continue;
            int32_t d = 0;
            if (ge_poll(slot.ctx, &d) != GE_OK || !d) continue;
            const uint32_t q = (uint32_t)slot.req;
            int rc = ge_sync(slot.ctx);
            slot.remaining -= slot.inflight;
            slot.inflight = 0;
            if (rc == GE_OK && slot.remaining > 0) continue;
            if (rc == GE_OK) rc = ge_download(slot.ctx, &reqs[q].out);
            finish(q, rc);
            slot.req = -1;
        }
    }
    // leave the batch empty for the next run
    if (!resident.empty()) (void)ge_systems_update(s->batch, 0, nullptr, 0, nullptr, nullptr);
    if (stats) {
        ge_server_stats st{};
        std::vector<double> lat;
        double first = 1e300, last = 0.0;
        for (uint32_t i = 0; i < nreq; ++i) {
            if (res[i].status != GE_OK) continue;
            lat.push_back(res[i].done_s - res[i].arrival_s);
            first = std::min(first, res[i].arrival_s);
            last = std::max(last, res[i].done_s);
            st.particle_updates += (uint64_t)reqs[i].n * (uint64_t)reqs[i].steps;
        }
        st.completed = (uint32_t)lat.size();
        if (!lat.empty()) {
            std::sort(lat.begin(), lat.end());
            auto pct = [&](double p) {
                const size_t k = (size_t)std::min<double>(lat.size() - 1, std::ceil(p * lat.size()) - 1);
                return lat[k];
            };
            st.p50_s = pct(0.50);
            st.p99_s = pct(0.99);
            double sum = 0.0;
            for (double v : lat) sum += v;
            st.mean_s = sum / lat.size();
            st.wall_s = last - std::max(0.0, first);
            if (st.wall_s > 0) {
                st.solves_per_s = st.completed / st.wall_s;
                st.updates_per_s = st.particle_updates / st.wall_s;
            }
        }
        st.max_resident = max_resident;
        st.device_bytes = s->device_bytes;
        st.rounds = rounds;
        *stats = st;
    }
    return GE_OK;
}

}  // extern "C"
