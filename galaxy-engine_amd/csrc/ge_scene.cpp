// ge_scene.cpp -- Galaxy Engine save files (version 190) to and from device
// state: the entry points of include/ge_scene.h.  Host code; the particles
// move through ge_upload / ge_download (and their 3D twins).
//
// The record layout restates SaveSystem's field-by-field writes
// (GE/src/UX/saveSystem.cpp:309-396) and reads (GE/include/UX/saveSystem.h:
// 228-342) with the field types of GE/include/Particles/particle.h: packed,
// no padding, vectors of 2 (2D) or 3 (3D) floats.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <new>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/ge_render.h"
#include "../../include/ge_scene.h"

namespace {

const char *const kSeparator = "---PARTICLE BINARY DATA---";
constexpr uint32_t kVersion = 190;  // SaveSystem::currentVersion (saveSystem.h:37-38)

// Field order of one record; 'V' is a glm::vec2 (2D) or glm::vec3 (3D).
enum Field : int {
    F_POS, F_PREDPOS, F_VEL, F_PREVVEL, F_PREDVEL, F_ACC, F_MASS, F_PRESS, F_PRESSTMP, F_PRESSF, F_DENS,
    F_PREDDENS, F_SPHMASS, F_RESTDENS, F_STIFF, F_VISC, F_COHESION, F_TEMP, F_KE, F_PREVKE, F_MORTON, F_ID,
    F_HOTPOINT, F_SOLIDIFIED,
    F_COLOR, F_PCOLOR, F_SCOLOR, F_SPHCOLOR, F_SIZE, F_UNIQUECOLOR, F_SOLID, F_SUBDIV, F_RESIZE, F_DARK, F_ISSPH,
    F_SELECTED, F_GRABBED, F_PREVSIZE, F_NEIGHBORS, F_TOTALRADIUS, F_LIFESPAN, F_SPHLABEL, F_PINNED, F_DRAWN,
    F_SPAWNITER, F_TURBULENCE, F_COUNT
};

// bytes of each field; 0 marks a vector
constexpr int kSize[F_COUNT] = {
    0, 0, 0, 0, 0, 0, 4, 4, 4, 0, 4,
    4, 4, 4, 4, 4, 4, 4, 4, 4, 8, 4,
    1, 1,
    4, 4, 4, 4, 4, 1, 1, 1, 1, 1, 1,
    1, 1, 4, 4, 4, 4, 4, 1, 1,
    4, 4};

struct Layout {
    int dim = 2;
    uint32_t off[F_COUNT] = {};
    uint32_t bytes = 0;
    explicit Layout(int d = 2) : dim(d) {
        uint32_t o = 0;
        for (int f = 0; f < F_COUNT; ++f) {
            off[f] = o;
            o += kSize[f] ? kSize[f] : 4u * d;
        }
        bytes = o;
    }
};

template <typename T>
T rd(const uint8_t *rec, uint32_t off) {
    T v;
    std::memcpy(&v, rec + off, sizeof v);
    return v;
}

template <typename T>
void wr(uint8_t *rec, uint32_t off, T v) {
    std::memcpy(rec + off, &v, sizeof v);
}

// A record with the values a spawned particle gets: ParticlePhysics(pos, vel,
// mass, ...) (particle.h:35-63) and ParticleRendering(color, size, ...)
// (particle.h:100-130) with white colour, size 1, lifeSpan -1.
void default_record(const Layout &L, uint8_t *rec) {
    std::memset(rec, 0, L.bytes);
    wr<float>(rec, L.off[F_TEMP], 288.0f);
    const uint8_t white[4] = {255, 255, 255, 255}, grey[4] = {128, 128, 128, 128};
    std::memcpy(rec + L.off[F_COLOR], white, 4);
    std::memcpy(rec + L.off[F_PCOLOR], white, 4);
    std::memcpy(rec + L.off[F_SCOLOR], white, 4);
    std::memcpy(rec + L.off[F_SPHCOLOR], grey, 4);
    wr<float>(rec, L.off[F_SIZE], 1.0f);
    wr<float>(rec, L.off[F_PREVSIZE], 1.0f);
    wr<float>(rec, L.off[F_LIFESPAN], -1.0f);
    wr<int32_t>(rec, L.off[F_SPAWNITER], 100000000);
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

// "1.5" with max_digits10, as yaml-cpp writes a float
std::string fmt_float(float v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.9g", (double)v);
    return buf;
}

}  // namespace

struct ge_scene {
    std::vector<std::pair<std::string, std::string>> kv;  // YAML documents in file order
    int is3d = 0;
    uint32_t version = kVersion;
    uint32_t gid = 0, gshape = 0, gwall = 0;
    uint32_t count = 0;
    std::vector<uint8_t> recs;
    std::vector<uint8_t> tail;
    mutable std::string err;  // the last failure, also of const calls
    Layout layout() const { return Layout(is3d ? 3 : 2); }
};

namespace {

// why the last ge_scene_load on this thread failed (no scene to hold it)
thread_local std::string g_load_error;

int fail(ge_scene *s, int code, const std::string &msg) {
    if (s) s->err = msg;
    return code;
}

const std::string *find_value(const ge_scene *s, const std::string &key) {
    for (const auto &p : s->kv)
        if (p.first == key) return &p.second;
    return nullptr;
}

void set_kv(ge_scene *s, const std::string &key, const std::string &value) {
    for (auto &p : s->kv)
        if (p.first == key) {
            p.second = value;
            return;
        }
    s->kv.emplace_back(key, value);
}

// yaml-cpp's scalar forms of a boolean (true/false, yes/no, on/off, y/n)
bool parse_bool(const std::string &v, bool &out) {
    std::string l;
    for (char ch : v) l += (char)std::tolower((unsigned char)ch);
    if (l == "true" || l == "yes" || l == "on" || l == "y") {
        out = true;
        return true;
    }
    if (l == "false" || l == "no" || l == "off" || l == "n") {
        out = false;
        return true;
    }
    return false;
}

bool get_float(const ge_scene *s, const char *key, float &out) {
    const std::string *v = find_value(s, key);
    if (!v) return false;
    char *end = nullptr;
    errno = 0;
    const float f = std::strtof(v->c_str(), &end);
    if (end == v->c_str() || errno == ERANGE) return false;
    out = f;
    return true;
}

bool get_bool(const ge_scene *s, const char *key, bool &out) {
    const std::string *v = find_value(s, key);
    return v && parse_bool(*v, out);
}

// The YAML header: one "Key: value" map per document (paramIO,
// saveSystem.h:40-61); lines of document separators and comments skipped.
void parse_yaml(ge_scene *s, const std::string &text) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#' || t.rfind("---", 0) == 0 || t == "...") continue;
        const size_t c = t.find(':');
        if (c == std::string::npos) continue;
        std::string key = trim(t.substr(0, c)), value = trim(t.substr(c + 1));
        if (value.size() >= 2 && ((value.front() == '"' && value.back() == '"') ||
                                  (value.front() == '\'' && value.back() == '\'')))
            value = value.substr(1, value.size() - 2);
        set_kv(s, key, value);
    }
}

std::string yaml_text(const ge_scene *s) {
    std::string out;
    bool first = true;
    for (const auto &p : s->kv) {
        if (!first) out += "---\n";
        out += p.first + ": " + p.second + "\n";
        first = false;
    }
    return out;
}

struct Soa {
    std::vector<float> x, y, z, vx, vy, vz, ax, ay, az, pvx, pvy, pvz, m;
    std::vector<uint32_t> flags, id;
    std::vector<float> sph_mass, rest_dens, stiff, visc, cohesion, dens, press, press_tmp, pred_dens;
    std::vector<float> ppx, ppy, ppz, pfx, pfy, pfz;
    std::vector<uint32_t> color, pcolor, scolor;  // render fields (RGBA8)
    std::vector<float> size, prev_size;
    std::vector<int32_t> nbrs;
    std::vector<float> temp, ke, prev_ke;          // temperature fields
    std::vector<uint32_t> label;
    std::vector<std::vector<float> *> floats() {
        return {&x, &y, &z, &vx, &vy, &vz, &ax, &ay, &az, &pvx, &pvy, &pvz, &m, &sph_mass, &rest_dens, &stiff,
                &visc, &cohesion, &dens, &press, &press_tmp, &pred_dens, &ppx, &ppy, &ppz, &pfx, &pfy, &pfz};
    }
    void resize(uint32_t n) {
        for (auto *v : floats()) v->assign(n, 0.0f);
        flags.assign(n, 0u);
        id.assign(n, 0u);
        color.assign(n, 0u);
        pcolor.assign(n, 0u);
        scolor.assign(n, 0u);
        size.assign(n, 0.0f);
        prev_size.assign(n, 0.0f);
        nbrs.assign(n, 0);
        temp.assign(n, 0.0f);
        ke.assign(n, 0.0f);
        prev_ke.assign(n, 0.0f);
        label.assign(n, 0u);
    }
    ge_thermal_fields thermal() {
        ge_thermal_fields f{};
        f.temp = temp.data();
        f.label = label.data();
        f.ke = ke.data();
        f.prev_ke = prev_ke.data();
        return f;
    }
    ge_render_fields render() {
        ge_render_fields f{};
        f.color = color.data();
        f.pcolor = pcolor.data();
        f.scolor = scolor.data();
        f.size = size.data();
        f.neighbors = nbrs.data();
        f.previous_size = prev_size.data();
        return f;
    }
};

void read_vec(const uint8_t *rec, uint32_t off, int dim, float &a, float &b, float &c) {
    a = rd<float>(rec, off);
    b = rd<float>(rec, off + 4);
    c = dim == 3 ? rd<float>(rec, off + 8) : 0.0f;
}

void write_vec(uint8_t *rec, uint32_t off, int dim, float a, float b, float c) {
    wr<float>(rec, off, a);
    wr<float>(rec, off + 4, b);
    if (dim == 3) wr<float>(rec, off + 8, c);
}

// records -> structure of arrays (the ParticlePhysics / ParticleRendering
// fields on the device path)
void decode(const ge_scene *s, Soa &o) {
    const Layout L = s->layout();
    const int d = L.dim;
    o.resize(s->count);
    for (uint32_t i = 0; i < s->count; ++i) {
        const uint8_t *r = s->recs.data() + (size_t)i * L.bytes;
        read_vec(r, L.off[F_POS], d, o.x[i], o.y[i], o.z[i]);
        read_vec(r, L.off[F_VEL], d, o.vx[i], o.vy[i], o.vz[i]);
        read_vec(r, L.off[F_ACC], d, o.ax[i], o.ay[i], o.az[i]);
        read_vec(r, L.off[F_PREVVEL], d, o.pvx[i], o.pvy[i], o.pvz[i]);
        read_vec(r, L.off[F_PREDPOS], d, o.ppx[i], o.ppy[i], o.ppz[i]);
        read_vec(r, L.off[F_PRESSF], d, o.pfx[i], o.pfy[i], o.pfz[i]);
        o.m[i] = rd<float>(r, L.off[F_MASS]);
        o.press[i] = rd<float>(r, L.off[F_PRESS]);
        o.press_tmp[i] = rd<float>(r, L.off[F_PRESSTMP]);
        o.dens[i] = rd<float>(r, L.off[F_DENS]);
        o.pred_dens[i] = rd<float>(r, L.off[F_PREDDENS]);
        o.sph_mass[i] = rd<float>(r, L.off[F_SPHMASS]);
        o.rest_dens[i] = rd<float>(r, L.off[F_RESTDENS]);
        o.stiff[i] = rd<float>(r, L.off[F_STIFF]);
        o.visc[i] = rd<float>(r, L.off[F_VISC]);
        o.cohesion[i] = rd<float>(r, L.off[F_COHESION]);
        o.id[i] = rd<uint32_t>(r, L.off[F_ID]);
        uint32_t f = 0;
        if (rd<uint8_t>(r, L.off[F_PINNED])) f |= GE_FLAG_PINNED;
        if (rd<uint8_t>(r, L.off[F_ISSPH])) f |= GE_FLAG_SPH;
        if (rd<uint8_t>(r, L.off[F_DRAWN])) f |= GE_FLAG_DRAWN;
        if (rd<uint8_t>(r, L.off[F_DARK])) f |= GE_FLAG_DARK;
        if (rd<uint8_t>(r, L.off[F_UNIQUECOLOR])) f |= GE_FLAG_UNIQUE_COLOR;
        if (rd<uint8_t>(r, L.off[F_SOLID])) f |= GE_FLAG_SOLID;
        if (rd<uint8_t>(r, L.off[F_HOTPOINT])) f |= GE_FLAG_HOT_POINT;
        if (rd<uint8_t>(r, L.off[F_SOLIDIFIED])) f |= GE_FLAG_SOLIDIFIED;
        o.flags[i] = f;
        o.temp[i] = rd<float>(r, L.off[F_TEMP]);
        o.ke[i] = rd<float>(r, L.off[F_KE]);
        o.prev_ke[i] = rd<float>(r, L.off[F_PREVKE]);
        o.label[i] = rd<uint32_t>(r, L.off[F_SPHLABEL]);
        o.color[i] = rd<uint32_t>(r, L.off[F_COLOR]);  // Color {r, g, b, a}: r in the low byte
        o.pcolor[i] = rd<uint32_t>(r, L.off[F_PCOLOR]);
        o.scolor[i] = rd<uint32_t>(r, L.off[F_SCOLOR]);
        o.size[i] = rd<float>(r, L.off[F_SIZE]);
        o.nbrs[i] = rd<int32_t>(r, L.off[F_NEIGHBORS]);
        o.prev_size[i] = rd<float>(r, L.off[F_PREVSIZE]);
    }
}

// A float the device never produces (a NaN with an all-ones payload): the
// download prefills the SPH arrays with it, so fields a context without SPH
// state leaves untouched keep the record's values.
constexpr uint32_t kUntouched = 0xFFFFFFFFu;
bool untouched(float v) {
    uint32_t u;
    std::memcpy(&u, &v, sizeof u);
    return u == kUntouched;
}

// structure of arrays (device order) -> records, carrying every field the
// device does not hold over from the record of the same id
void encode(ge_scene *s, const Soa &o, uint32_t n, bool render, bool thermal) {
    const Layout L = s->layout();
    const int d = L.dim;
    std::unordered_map<uint32_t, uint32_t> by_id;
    by_id.reserve(s->count * 2 + 1);
    for (uint32_t i = 0; i < s->count; ++i)
        by_id.emplace(rd<uint32_t>(s->recs.data() + (size_t)i * L.bytes, L.off[F_ID]), i);
    std::vector<uint8_t> out((size_t)n * L.bytes);
    uint32_t max_id = 0;
    for (uint32_t k = 0; k < n; ++k) {
        uint8_t *r = out.data() + (size_t)k * L.bytes;
        const auto it = by_id.find(o.id[k]);
        if (it != by_id.end()) std::memcpy(r, s->recs.data() + (size_t)it->second * L.bytes, L.bytes);
        else default_record(L, r);
        auto keep = [&](uint32_t off, float v) {
            if (!untouched(v)) wr<float>(r, off, v);
        };
        write_vec(r, L.off[F_POS], d, o.x[k], o.y[k], o.z[k]);
        write_vec(r, L.off[F_VEL], d, o.vx[k], o.vy[k], o.vz[k]);
        write_vec(r, L.off[F_ACC], d, o.ax[k], o.ay[k], o.az[k]);
        write_vec(r, L.off[F_PREVVEL], d, o.pvx[k], o.pvy[k], o.pvz[k]);
        wr<float>(r, L.off[F_MASS], o.m[k]);
        keep(L.off[F_PREDPOS], o.ppx[k]);
        keep(L.off[F_PREDPOS] + 4, o.ppy[k]);
        keep(L.off[F_PRESSF], o.pfx[k]);
        keep(L.off[F_PRESSF] + 4, o.pfy[k]);
        if (d == 3) {
            keep(L.off[F_PREDPOS] + 8, o.ppz[k]);
            keep(L.off[F_PRESSF] + 8, o.pfz[k]);
        }
        keep(L.off[F_PRESS], o.press[k]);
        keep(L.off[F_PRESSTMP], o.press_tmp[k]);
        keep(L.off[F_DENS], o.dens[k]);
        keep(L.off[F_PREDDENS], o.pred_dens[k]);
        keep(L.off[F_SPHMASS], o.sph_mass[k]);
        keep(L.off[F_RESTDENS], o.rest_dens[k]);
        keep(L.off[F_STIFF], o.stiff[k]);
        keep(L.off[F_VISC], o.visc[k]);
        keep(L.off[F_COHESION], o.cohesion[k]);
        wr<uint32_t>(r, L.off[F_ID], o.id[k]);
        wr<uint8_t>(r, L.off[F_PINNED], (o.flags[k] & GE_FLAG_PINNED) ? 1 : 0);
        wr<uint8_t>(r, L.off[F_ISSPH], (o.flags[k] & GE_FLAG_SPH) ? 1 : 0);
        wr<uint8_t>(r, L.off[F_DRAWN], (o.flags[k] & GE_FLAG_DRAWN) ? 1 : 0);
        wr<uint8_t>(r, L.off[F_DARK], (o.flags[k] & GE_FLAG_DARK) ? 1 : 0);
        wr<uint8_t>(r, L.off[F_UNIQUECOLOR], (o.flags[k] & GE_FLAG_UNIQUE_COLOR) ? 1 : 0);
        wr<uint8_t>(r, L.off[F_SOLID], (o.flags[k] & GE_FLAG_SOLID) ? 1 : 0);
        wr<uint8_t>(r, L.off[F_HOTPOINT], (o.flags[k] & GE_FLAG_HOT_POINT) ? 1 : 0);
        wr<uint8_t>(r, L.off[F_SOLIDIFIED], (o.flags[k] & GE_FLAG_SOLIDIFIED) ? 1 : 0);
        if (thermal) {
            wr<float>(r, L.off[F_TEMP], o.temp[k]);
            wr<float>(r, L.off[F_KE], o.ke[k]);
            wr<float>(r, L.off[F_PREVKE], o.prev_ke[k]);
            wr<uint32_t>(r, L.off[F_SPHLABEL], o.label[k]);
        }
        if (render) {
            wr<uint32_t>(r, L.off[F_COLOR], o.color[k]);
            wr<uint32_t>(r, L.off[F_PCOLOR], o.pcolor[k]);
            wr<uint32_t>(r, L.off[F_SCOLOR], o.scolor[k]);
            wr<float>(r, L.off[F_SIZE], o.size[k]);
            wr<int32_t>(r, L.off[F_NEIGHBORS], o.nbrs[k]);
            wr<float>(r, L.off[F_PREVSIZE], o.prev_size[k]);
        }
        if (o.id[k] > max_id) max_id = o.id[k];
    }
    s->recs.swap(out);
    s->count = n;
    if (n && max_id + 1 > s->gid) s->gid = max_id + 1;  // globalId = maxId + 1 (saveSystem.h:191-196)
}

// the constraints at the start of the tail: uint32 count, then count raw
// ParticleConstraint records (saveSystem.cpp:421-429)
bool tail_constraints(const ge_scene *s, std::vector<ge_constraint> &out) {
    out.clear();
    if (s->tail.size() < 4) return true;  // an empty tail: no constraints
    const uint32_t m = rd<uint32_t>(s->tail.data(), 0);
    if (s->tail.size() < 4 + (uint64_t)m * sizeof(ge_constraint)) return false;
    out.resize(m);
    if (m) std::memcpy(out.data(), s->tail.data() + 4, (size_t)m * sizeof(ge_constraint));
    return true;
}

bool set_tail_constraints(ge_scene *s, const std::vector<ge_constraint> &cons) {
    uint32_t old = 0;
    if (s->tail.size() >= 4) {
        old = rd<uint32_t>(s->tail.data(), 0);
        if (s->tail.size() < 4 + (uint64_t)old * sizeof(ge_constraint)) return false;
    }
    std::vector<uint8_t> t(4 + cons.size() * sizeof(ge_constraint));
    wr<uint32_t>(t.data(), 0, (uint32_t)cons.size());
    if (!cons.empty()) std::memcpy(t.data() + 4, cons.data(), cons.size() * sizeof(ge_constraint));
    const size_t rest = s->tail.size() >= 4 ? 4 + (size_t)old * sizeof(ge_constraint) : 0;
    t.insert(t.end(), s->tail.begin() + (std::ptrdiff_t)rest, s->tail.end());
    s->tail.swap(t);
    return true;
}

template <typename P>
P view_of(Soa &o) {
    P p{};
    p.x = o.x.data(); p.y = o.y.data(); p.vx = o.vx.data(); p.vy = o.vy.data();
    p.ax = o.ax.data(); p.ay = o.ay.data(); p.pvx = o.pvx.data(); p.pvy = o.pvy.data();
    p.mass = o.m.data(); p.flags = o.flags.data(); p.id = o.id.data();
    p.sph_mass = o.sph_mass.data(); p.rest_dens = o.rest_dens.data(); p.stiff = o.stiff.data();
    p.visc = o.visc.data(); p.cohesion = o.cohesion.data(); p.dens = o.dens.data(); p.press = o.press.data();
    p.press_tmp = o.press_tmp.data(); p.pred_dens = o.pred_dens.data(); p.ppx = o.ppx.data(); p.ppy = o.ppy.data();
    p.pfx = o.pfx.data(); p.pfy = o.pfy.data();
    return p;
}

void prefill_untouched(Soa &o) {
    float u;
    std::memcpy(&u, &kUntouched, sizeof u);
    for (auto *v : {&o.sph_mass, &o.rest_dens, &o.stiff, &o.visc, &o.cohesion, &o.dens, &o.press, &o.press_tmp,
                    &o.pred_dens, &o.ppx, &o.ppy, &o.ppz, &o.pfx, &o.pfy, &o.pfz})
        v->assign(v->size(), u);
}

}  // namespace

extern "C" {

int ge_scene_new(int32_t is3d, ge_scene **out) {
    if (!out) return GE_E_INVALID;
    ge_scene *s = new (std::nothrow) ge_scene();
    if (!s) return GE_E_NOMEM;
    s->is3d = is3d ? 1 : 0;
    s->gid = 0;
    set_kv(s, "is3DMode", s->is3d ? "true" : "false");
    s->tail.assign(6 * sizeof(uint32_t), 0);  // no constraints, walls, shapes, point, area or cone lights
    *out = s;
    return GE_OK;
}

int ge_scene_load(const char *path, ge_scene **out) {
    if (!path || !out) return GE_E_INVALID;
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        g_load_error = std::string("cannot open ") + path;
        return GE_E_INVALID;
    }
    std::vector<uint8_t> all((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    ge_scene *s = new (std::nothrow) ge_scene();
    if (!s) return GE_E_NOMEM;
    // the YAML text ends at the line holding the separator token
    // (deserializeParticleSystem, saveSystem.h:136-149)
    size_t pos = 0, bin = std::string::npos;
    std::string yaml;
    const std::string tok(kSeparator);
    while (pos < all.size()) {
        size_t e = pos;
        while (e < all.size() && all[e] != '\n') ++e;
        const std::string line(all.begin() + pos, all.begin() + e);
        if (line.find(tok) != std::string::npos) {
            bin = e < all.size() ? e + 1 : e;
            break;
        }
        yaml += line + "\n";
        pos = e + 1;
    }
    auto bad = [&](int code, const char *msg) {
        g_load_error = msg;
        delete s;
        return code;
    };
    if (bin == std::string::npos) return bad(GE_E_INVALID, "separator not found");
    parse_yaml(s, yaml);
    bool is3d = false;
    get_bool(s, "is3DMode", is3d);
    s->is3d = is3d ? 1 : 0;
    const size_t need = 5 * sizeof(uint32_t);
    if (all.size() - bin < need) return bad(GE_E_INVALID, "binary section truncated");
    const uint8_t *b = all.data() + bin;
    s->version = rd<uint32_t>(b, 0);
    if (s->version != 190 && s->version != 180)  // the versions deserializeParticleSystem reads
        return bad(GE_E_UNSUPPORTED, "unsupported save version");
    s->gid = rd<uint32_t>(b, 4);
    s->gshape = rd<uint32_t>(b, 8);
    s->gwall = rd<uint32_t>(b, 12);
    s->count = rd<uint32_t>(b, 16);
    const Layout L = s->layout();
    const uint64_t rec_bytes = (uint64_t)s->count * L.bytes;
    if ((uint64_t)(all.size() - bin - need) < rec_bytes) return bad(GE_E_INVALID, "particle records truncated");
    s->recs.assign(b + need, b + need + rec_bytes);
    s->tail.assign(b + need + rec_bytes, static_cast<const uint8_t *>(all.data() + all.size()));
    *out = s;
    return GE_OK;
}

void ge_scene_free(ge_scene *s) { delete s; }

const char *ge_scene_last_error(const ge_scene *s) { return s ? s->err.c_str() : g_load_error.c_str(); }

int ge_scene_get_info(const ge_scene *s, ge_scene_info *info) {
    if (!s || !info) return GE_E_INVALID;
    std::memset(info, 0, sizeof *info);
    info->is3d = s->is3d;
    info->version = s->version;
    info->count = s->count;
    info->record_bytes = s->layout().bytes;
    info->global_id = s->gid;
    info->global_shape_id = s->gshape;
    info->global_wall_id = s->gwall;
    info->tail_bytes = s->tail.size();
    return GE_OK;
}

const char *ge_scene_value(const ge_scene *s, const char *key) {
    if (!s || !key) return nullptr;
    const std::string *v = find_value(s, key);
    return v ? v->c_str() : nullptr;
}

int ge_scene_set_value(ge_scene *s, const char *key, const char *value) {
    if (!s || !key || !value || !*key) return GE_E_INVALID;
    if (std::strchr(key, '\n') || std::strchr(value, '\n') || std::strchr(key, ':'))
        return fail(s, GE_E_INVALID, "keys and values are single-line, keys have no ':'");
    set_kv(s, key, value);
    if (std::strcmp(key, "is3DMode") == 0) {
        bool b = false;
        if (!parse_bool(value, b)) return fail(s, GE_E_INVALID, "is3DMode must be a boolean");
        if ((b ? 1 : 0) != s->is3d && s->count)
            return fail(s, GE_E_INVALID, "is3DMode cannot change while the scene holds particles");
        s->is3d = b ? 1 : 0;
    }
    return GE_OK;
}

int ge_scene_records(const ge_scene *s, const void **data, uint64_t *bytes) {
    if (!s || !data || !bytes) return GE_E_INVALID;
    *data = s->recs.data();
    *bytes = s->recs.size();
    return GE_OK;
}

int ge_scene_params(const ge_scene *s, ge_step_params *p) {
    if (!s || !p) return GE_E_INVALID;
    ge_default_params(p);
    float v;
    bool b, inf = false;
    if (get_float(s, "GravityMultiplier", v)) p->G = (float)(6.674e-11 * (double)v);  // globalLogic.cpp:1135
    if (get_float(s, "TimeMult", v)) p->dt = p->dt_start = 0.045f * v;                // globalLogic.cpp:1187
    if (get_float(s, "Theta", v)) p->theta = v;
    if (get_float(s, "Softening", v)) p->softening = v;
    if (get_bool(s, "SPHEnabled", b)) p->sph = b;
    if (get_float(s, "SPHRadiusMult", v)) p->sph_h = v;
    if (get_float(s, "SPHMass", v)) p->sph_unit_mass = v;
    if (get_float(s, "SPHViscosity", v)) p->viscosity = v;
    if (get_float(s, "SPHStiffness", v)) p->stiff_mult = v;
    if (get_float(s, "SPHCohesion", v)) p->cohesion_coef = v;
    if (get_float(s, "SPHDelta", v)) p->delta = v;
    if (get_float(s, "SPHMaxVel", v)) p->sph_max_vel = v;
    if (get_bool(s, "SPHGround", b)) p->sph_ground = b;
    if (get_float(s, "SPHGravity", v)) p->vertical_gravity = v;
    get_bool(s, "InfiniteDomain", inf);
    if (get_bool(s, "LoopingSpace", b)) p->periodic = b && !inf;
    if (get_float(s, "DomainWidth", v)) p->domain_w = v;
    if (get_float(s, "DomainHeight", v)) p->domain_h = v;
    if (get_bool(s, "TemperatureSimulation", b)) p->temperature = b;
    return GE_OK;
}

int ge_scene_params3(const ge_scene *s, ge3_step_params *p) {
    if (!s || !p) return GE_E_INVALID;
    ge3_default_params(p);
    float v;
    bool b, inf = false;
    if (get_float(s, "GravityMultiplier", v)) p->G = (float)(6.674e-11 * (double)v);
    if (get_float(s, "TimeMult", v)) p->dt = p->dt_start = 0.045f * v;
    if (get_float(s, "Theta", v)) p->theta = v;
    if (get_float(s, "Softening", v)) p->softening = v;
    if (get_bool(s, "SPHEnabled", b)) p->sph = b;
    if (get_float(s, "SPHRadiusMult", v)) p->sph_h = v;
    if (get_float(s, "SPHMass", v)) p->sph_unit_mass = v;
    if (get_float(s, "SPHViscosity", v)) p->viscosity = v;
    if (get_float(s, "SPHStiffness", v)) p->stiff_mult = v;
    if (get_float(s, "SPHCohesion", v)) p->cohesion_coef = v;
    if (get_float(s, "SPHDelta", v)) p->delta = v;
    if (get_float(s, "SPHMaxVel", v)) p->sph_max_vel = v;
    if (get_bool(s, "SPHGround", b)) p->sph_ground = b;
    if (get_float(s, "SPHGravity", v)) p->vertical_gravity = v;
    get_bool(s, "InfiniteDomain", inf);
    if (get_bool(s, "LoopingSpace", b)) p->periodic = b && !inf;
    if (get_float(s, "DomainWidth3D", v)) p->domain_w = v;
    if (get_float(s, "DomainHeight3D", v)) p->domain_h = v;
    if (get_float(s, "DomainDepth3D", v)) p->domain_d = v;
    if (get_bool(s, "TemperatureSimulation", b)) p->temperature = b;
    return GE_OK;
}

int ge_scene_set_params(ge_scene *s, const ge_step_params *p) {
    if (!s || !p) return GE_E_INVALID;
    auto B = [](int32_t v) { return std::string(v ? "true" : "false"); };
    set_kv(s, "GravityMultiplier", fmt_float((float)((double)p->G / 6.674e-11)));
    set_kv(s, "TimeMult", fmt_float(p->dt / 0.045f));
    set_kv(s, "Theta", fmt_float(p->theta));
    set_kv(s, "Softening", fmt_float(p->softening));
    set_kv(s, "SPHEnabled", B(p->sph));
    set_kv(s, "SPHRadiusMult", fmt_float(p->sph_h));
    set_kv(s, "SPHMass", fmt_float(p->sph_unit_mass));
    set_kv(s, "SPHViscosity", fmt_float(p->viscosity));
    set_kv(s, "SPHStiffness", fmt_float(p->stiff_mult));
    set_kv(s, "SPHCohesion", fmt_float(p->cohesion_coef));
    set_kv(s, "SPHDelta", fmt_float(p->delta));
    set_kv(s, "SPHMaxVel", fmt_float(p->sph_max_vel));
    set_kv(s, "SPHGround", B(p->sph_ground));
    set_kv(s, "SPHGravity", fmt_float(p->vertical_gravity));
    set_kv(s, "LoopingSpace", B(p->periodic));
    set_kv(s, "InfiniteDomain", B(!p->periodic));
    set_kv(s, "DomainWidth", fmt_float(p->domain_w));
    set_kv(s, "DomainHeight", fmt_float(p->domain_h));
    set_kv(s, "TemperatureSimulation", B(p->temperature));
    return GE_OK;
}

int ge_scene_set_params3(ge_scene *s, const ge3_step_params *p) {
    if (!s || !p) return GE_E_INVALID;
    auto B = [](int32_t v) { return std::string(v ? "true" : "false"); };
    set_kv(s, "GravityMultiplier", fmt_float((float)((double)p->G / 6.674e-11)));
    set_kv(s, "TimeMult", fmt_float(p->dt / 0.045f));
    set_kv(s, "Theta", fmt_float(p->theta));
    set_kv(s, "Softening", fmt_float(p->softening));
    set_kv(s, "SPHEnabled", B(p->sph));
    set_kv(s, "SPHRadiusMult", fmt_float(p->sph_h));
    set_kv(s, "SPHMass", fmt_float(p->sph_unit_mass));
    set_kv(s, "SPHViscosity", fmt_float(p->viscosity));
    set_kv(s, "SPHStiffness", fmt_float(p->stiff_mult));
    set_kv(s, "SPHCohesion", fmt_float(p->cohesion_coef));
    set_kv(s, "SPHDelta", fmt_float(p->delta));
    set_kv(s, "SPHMaxVel", fmt_float(p->sph_max_vel));
    set_kv(s, "SPHGround", B(p->sph_ground));
    set_kv(s, "SPHGravity", fmt_float(p->vertical_gravity));
    set_kv(s, "LoopingSpace", B(p->periodic));
    set_kv(s, "InfiniteDomain", B(!p->periodic));
    set_kv(s, "DomainWidth3D", fmt_float(p->domain_w));
    set_kv(s, "DomainHeight3D", fmt_float(p->domain_h));
    set_kv(s, "DomainDepth3D", fmt_float(p->domain_d));
    set_kv(s, "TemperatureSimulation", B(p->temperature));
    return GE_OK;
}

int ge_scene_constraint_params(const ge_scene *s, ge_constraint_params *c, int32_t *enabled) {
    if (!s || !c) return GE_E_INVALID;
    ge_default_constraint_params(c);
    bool b = false;
    if (enabled) *enabled = get_bool(s, "ParticleConstraints", b) && b ? 1 : 0;
    if (get_bool(s, "UnbreakableConstraints", b)) c->unbreakable = b ? 1 : 0;
    get_float(s, "ConstraintsStiffMultiplier", c->stiffness_mult);
    get_float(s, "ConstraintsResistMultiplier", c->resistance_mult);
    return GE_OK;
}

int ge_scene_set_constraint_params(ge_scene *s, const ge_constraint_params *c, int32_t enabled) {
    if (!s || !c) return GE_E_INVALID;
    set_kv(s, "ParticleConstraints", enabled ? "true" : "false");
    set_kv(s, "UnbreakableConstraints", c->unbreakable ? "true" : "false");
    set_kv(s, "ConstraintsStiffMultiplier", fmt_float(c->stiffness_mult));
    set_kv(s, "ConstraintsResistMultiplier", fmt_float(c->resistance_mult));
    return GE_OK;
}

int ge_scene_thermal_params(const ge_scene *s, ge_thermal_params *t) {
    if (!s || !t) return GE_E_INVALID;
    ge_default_thermal_params(t);
    get_float(s, "AmbientTemperature", t->ambient_temp);
    get_float(s, "AmbientHeatRate", t->ambient_heat_rate);
    get_float(s, "HeatConductivityMultiplier", t->heat_conductivity);
    return GE_OK;
}

int ge_scene_set_thermal_params(ge_scene *s, const ge_thermal_params *t) {
    if (!s || !t) return GE_E_INVALID;
    set_kv(s, "AmbientTemperature", fmt_float(t->ambient_temp));
    set_kv(s, "AmbientHeatRate", fmt_float(t->ambient_heat_rate));
    set_kv(s, "HeatConductivityMultiplier", fmt_float(t->heat_conductivity));
    return GE_OK;
}

int ge_scene_upload(const ge_scene *s, ge_ctx *ctx) {
    if (!s || !ctx) return GE_E_INVALID;
    if (s->is3d) {
        s->err = "a 3D scene takes a 3D context";
        return GE_E_INVALID;
    }
    Soa o;
    decode(s, o);
    ge_particles p = view_of<ge_particles>(o);
    int r = ge_upload(ctx, s->count, &p);
    int32_t render = 0;
    if (r == GE_OK && (r = ge_render_enabled(ctx, &render)) == GE_OK && render) {
        ge_render_fields f = o.render();
        r = ge_render_upload(ctx, &f);
    }
    int32_t thermal = 0;
    if (r == GE_OK && (r = ge_thermal_enabled(ctx, &thermal)) == GE_OK && thermal) {
        ge_thermal_fields f = o.thermal();
        r = ge_thermal_upload(ctx, &f);
    }
    uint32_t ccap = 0;
    if (r == GE_OK && (r = ge_constraints_capacity(ctx, &ccap)) == GE_OK && ccap) {
        std::vector<ge_constraint> cons;
        if (!tail_constraints(s, cons)) {
            s->err = "constraint section of the tail truncated";
            return GE_E_INVALID;
        }
        r = ge_constraints_upload(ctx, cons.data(), (uint32_t)cons.size());
    }
    if (r != GE_OK) s->err = ge_last_error(ctx);
    return r;
}

int ge_scene_upload3(const ge_scene *s, ge3_ctx *ctx) {
    if (!s || !ctx) return GE_E_INVALID;
    if (!s->is3d) {
        s->err = "a 2D scene takes a 2D context";
        return GE_E_INVALID;
    }
    Soa o;
    decode(s, o);
    ge3_particles p = view_of<ge3_particles>(o);
    p.z = o.z.data(); p.vz = o.vz.data(); p.az = o.az.data(); p.pvz = o.pvz.data();
    p.ppz = o.ppz.data(); p.pfz = o.pfz.data();
    int r = ge3_upload(ctx, s->count, &p);
    int32_t render = 0;
    if (r == GE_OK && (r = ge3_render_enabled(ctx, &render)) == GE_OK && render) {
        ge_render_fields f = o.render();
        r = ge3_render_upload(ctx, &f);
    }
    if (r != GE_OK) s->err = ge3_last_error(ctx);
    return r;
}

int ge_scene_download(ge_scene *s, ge_ctx *ctx) {
    if (!s || !ctx) return GE_E_INVALID;
    if (s->is3d) return fail(s, GE_E_INVALID, "a 3D scene takes a 3D context");
    uint32_t n = 0;
    if (int r = ge_count(ctx, &n)) return r;
    Soa o;
    o.resize(n);
    prefill_untouched(o);
    ge_particles p = view_of<ge_particles>(o);
    if (int r = ge_download(ctx, &p)) return fail(s, r, ge_last_error(ctx));
    int32_t render = 0;
    if (int r = ge_render_enabled(ctx, &render)) return fail(s, r, ge_last_error(ctx));
    if (render) {
        ge_render_fields f = o.render();
        if (int r = ge_render_download(ctx, &f)) return fail(s, r, ge_last_error(ctx));
    }
    int32_t thermal = 0;
    if (int r = ge_thermal_enabled(ctx, &thermal)) return fail(s, r, ge_last_error(ctx));
    if (thermal) {
        ge_thermal_fields f = o.thermal();
        if (int r = ge_thermal_download(ctx, &f)) return fail(s, r, ge_last_error(ctx));
    }
    uint32_t ccap = 0;
    if (int r = ge_constraints_capacity(ctx, &ccap)) return fail(s, r, ge_last_error(ctx));
    if (ccap) {
        uint32_t m = 0;
        if (int r = ge_constraints_download(ctx, nullptr, 0, &m)) return fail(s, r, ge_last_error(ctx));
        std::vector<ge_constraint> cons(m);
        if (m) {
            if (int r = ge_constraints_download(ctx, cons.data(), m, &m)) return fail(s, r, ge_last_error(ctx));
        }
        if (!set_tail_constraints(s, cons)) return fail(s, GE_E_INVALID, "constraint section of the tail truncated");
    }
    encode(s, o, n, render != 0, thermal != 0);
    return GE_OK;
}

int ge_scene_download3(ge_scene *s, ge3_ctx *ctx) {
    if (!s || !ctx) return GE_E_INVALID;
    if (!s->is3d) return fail(s, GE_E_INVALID, "a 2D scene takes a 2D context");
    uint32_t n = 0;
    if (int r = ge3_count(ctx, &n)) return r;
    Soa o;
    o.resize(n);
    prefill_untouched(o);
    ge3_particles p = view_of<ge3_particles>(o);
    p.z = o.z.data(); p.vz = o.vz.data(); p.az = o.az.data(); p.pvz = o.pvz.data();
    p.ppz = o.ppz.data(); p.pfz = o.pfz.data();
    if (int r = ge3_download(ctx, &p)) return fail(s, r, ge3_last_error(ctx));
    int32_t render = 0;
    if (int r = ge3_render_enabled(ctx, &render)) return fail(s, r, ge3_last_error(ctx));
    if (render) {
        ge_render_fields f = o.render();
        if (int r = ge3_render_download(ctx, &f)) return fail(s, r, ge3_last_error(ctx));
    }
    encode(s, o, n, render != 0, false);
    return GE_OK;
}

int ge_scene_save(const ge_scene *s, const char *path) {
    if (!s || !path) return GE_E_INVALID;
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return GE_E_INVALID;
    const std::string yaml = yaml_text(s);
    f.write(yaml.data(), (std::streamsize)yaml.size());
    const std::string sep = std::string("\n") + kSeparator + "\n";  // saveSystem.cpp:297
    f.write(sep.data(), (std::streamsize)sep.size());
    const uint32_t head[5] = {kVersion, s->gid, s->gshape, s->gwall, s->count};
    f.write(reinterpret_cast<const char *>(head), sizeof head);
    f.write(reinterpret_cast<const char *>(s->recs.data()), (std::streamsize)s->recs.size());
    f.write(reinterpret_cast<const char *>(s->tail.data()), (std::streamsize)s->tail.size());
    return f.good() ? GE_OK : GE_E_INVALID;
}

}  // extern "C"
