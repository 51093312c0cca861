/*
 * ge_oracle.c -- CPU restatement of Galaxy Engine's 2D physics step.
 *
 * TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline); the product
 * library never links it.  Build: oracle/Makefile (-O2 -mavx2 -mfma
 * -ffp-contract=off -fopenmp, the reference's flags from CMakeLists.txt:56-62
 * with contraction pinned off).
 *
 * Parity pins (what checks each function independently of this file) are
 * listed in tests/test_oracle.py; functions and the reference lines they
 * follow are cited inline.  GE/ = GalaxyEngine/.
 */
#include "ge_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define U32MAX 0xFFFFFFFFu

/* glm::min / glm::max scalar forms (glm 1.0.1 func_common.inl) */
static inline float fmin_glm(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_glm(float a, float b) { return (a < b) ? b : a; }

/* ---------------------------------------------------------------- bbox --- */
/* GE/src/globalLogic.cpp:1083-1099 */
void geo_bbox(int32_t n, const float *x, const float *y, float out[3]) {
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    for (int32_t i = 0; i < n; ++i) {
        mnx = fmin_glm(mnx, x[i]);
        mny = fmin_glm(mny, y[i]);
        mxx = fmax_glm(mxx, x[i]);
        mxy = fmax_glm(mxy, y[i]);
    }
    float size = fmax_glm(mxx - mnx, mxy - mny);
    float cx = (mnx + mxx) * 0.5f;
    float cy = (mny + mxy) * 0.5f;
    float half = size * 0.5f;
    out[0] = cx - half;
    out[1] = cy - half;
    out[2] = size;
}

/* -------------------------------------------------------------- Morton --- */
/* GE/src/Physics/morton.cpp:3-10 */
uint64_t geo_scale_to_grid(float pos, float lo, float hi) {
    if (hi <= lo) return 0;
    float c = (pos < lo) ? lo : ((hi < pos) ? hi : pos); /* std::clamp */
    float norm = (c - lo) / (hi - lo);
    uint64_t s = (uint64_t)(norm * 262144.0f);
    return s < 262143u ? s : 262143u;
}

/* GE/src/Physics/morton.cpp:12-20: 18-bit magic-mask spread */
uint64_t geo_spread_bits(uint64_t v) {
    v &= 0x3FFFFull;
    v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
    v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

/* GE/src/Physics/morton.cpp:22-25: x in even bits, y in odd bits */
uint64_t geo_morton2d(uint64_t ix, uint64_t iy) {
    return geo_spread_bits(ix) | (geo_spread_bits(iy) << 1);
}

/* GE/src/Physics/morton.cpp:27-38 */
void geo_morton_keys(int32_t n, const float *x, const float *y, const float bb[3], uint64_t *keys) {
    float ext = bb[2] < 1e-6f ? 1e-6f : bb[2]; /* std::max(posSize.z, 1e-6f) */
    float maxx = bb[0] + ext;
    float maxy = bb[1] + ext;
    for (int32_t i = 0; i < n; ++i) {
        uint64_t ix = geo_scale_to_grid(x[i], bb[0], maxx);
        uint64_t iy = geo_scale_to_grid(y[i], bb[1], maxy);
        keys[i] = geo_morton2d(ix, iy);
    }
}

/* stable sort by (key, index): LSD radix over 4 x 9-bit digits of the 36-bit key */
void geo_sort_perm(int32_t n, const uint64_t *keys, uint32_t *perm) {
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t *a = perm, *b = tmp;
    for (int32_t i = 0; i < n; ++i) a[i] = (uint32_t)i;
    /* keys may use more than 36 bits for generic callers: sort all 64 bits */
    for (int shift = 0; shift < 64; shift += 8) {
        uint32_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        int any = 0;
        for (int32_t i = 0; i < n; ++i) {
            uint32_t d = (uint32_t)(keys[a[i]] >> shift) & 0xFFu;
            cnt[d + 1]++;
            any |= (int)(d != 0);
        }
        if (!any) continue;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (int32_t i = 0; i < n; ++i) {
            uint32_t d = (uint32_t)(keys[a[i]] >> shift) & 0xFFu;
            b[cnt[d]++] = a[i];
        }
        uint32_t *t = a; a = b; b = t;
    }
    if (a != perm) memcpy(perm, a, sizeof(uint32_t) * (size_t)n);
    free(tmp);
}

/* ------------------------------------------------------------ quadtree --- */
typedef struct {
    geo_node *nodes;
    int64_t count, cap;
    const float *x, *y, *m;
    uint32_t leaf_cap;
} tree_builder;

/* std::partition_point over [first,last) with predicate v[i] < mid: the
 * libstdc++ halving search (quadtree.cpp:42-65 calls it three times). */
static uint32_t part_point(const float *v, uint32_t first, uint32_t last, float mid) {
    uint32_t len = last - first;
    while (len > 0) {
        uint32_t half = len >> 1;
        uint32_t middle = first + half;
        if (v[middle] < mid) {
            first = middle + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return first;
}

static int64_t tb_alloc(tree_builder *t) {
    if (t->count == t->cap) {
        int64_t nc = t->cap * 2 + 16;
        geo_node *nn = (geo_node *)realloc(t->nodes, sizeof(geo_node) * (size_t)nc);
        if (!nn) return -1;
        t->nodes = nn;
        t->cap = nc;
    }
    return t->count++;
}

/* Node::Node (GE/src/Physics/quadtree.cpp:5-26), subGridMaker (:28-88),
 * computeLeafMass / computeInternalMass / calculateNextNeighbor
 * (GE/include/Physics/quadtree.h:53-111). */
static int tb_node(tree_builder *t, int64_t idx, float px, float py, float size,
                   uint32_t s, uint32_t e) {
    geo_node nd;
    nd.posx = px; nd.posy = py; nd.size = size;
    nd.start = s; nd.end = e; nd.next = 0;
    nd.mass = 0.0f; nd.comx = 0.0f; nd.comy = 0.0f;
    nd.child[0] = nd.child[1] = nd.child[2] = nd.child[3] = U32MAX;
    uint32_t cnt = e - s;
    /* leaf_cap 0: the reference rule (quadtree.cpp:16-17); otherwise a node of
     * at most leaf_cap particles is a leaf (autotuned trees, same size floor) */
    int is_leaf = t->leaf_cap ? (cnt <= t->leaf_cap || size <= 0.000001f)
                              : (((cnt <= 4 && size <= 2.0f) || cnt == 1) || size <= 0.000001f);
    if (is_leaf) {
        /* computeLeafMass, quadtree.h:53-67 */
        float gm = 0.0f, cx = 0.0f, cy = 0.0f;
        for (uint32_t i = s; i < e; ++i) {
            gm = gm + t->m[i];
            cx = cx + t->x[i] * t->m[i];
            cy = cy + t->y[i] * t->m[i];
        }
        if (gm > 0) { cx = cx / gm; cy = cy / gm; }
        nd.mass = gm; nd.comx = cx; nd.comy = cy;
        t->nodes[idx] = nd;
        return 0;
    }
    /* subGridMaker, quadtree.cpp:28-88 */
    float half = size * 0.5f;
    float midx = px + half, midy = py + half;
    uint32_t b[5];
    b[0] = s; b[4] = e;
    b[2] = part_point(t->y, b[0], b[4], midy);
    b[1] = part_point(t->x, b[0], b[2], midx);
    b[3] = part_point(t->x, b[2], b[4], midx);
    t->nodes[idx] = nd; /* parent slot holds a placeholder while children build */
    for (int q = 0; q < 4; ++q) {
        if (b[q + 1] > b[q]) {
            float nx = px + ((q & 1) ? half : 0.0f);
            float ny = py + ((q & 2) ? half : 0.0f);
            int64_t ci = tb_alloc(t);
            if (ci < 0) return -1;
            if (tb_node(t, ci, nx, ny, half, b[q], b[q + 1]) != 0) return -1;
            nd.child[q] = (uint32_t)ci;
        }
    }
    /* computeInternalMass: loops subGrids[lr][ud], i.e. q = 0, 2, 1, 3 */
    static const int order[4] = {0, 2, 1, 3};
    float gm = 0.0f, cx = 0.0f, cy = 0.0f;
    uint32_t next = 0;
    for (int k = 0; k < 4; ++k) {
        uint32_t c = nd.child[order[k]];
        if (c == U32MAX) continue;
        const geo_node *ch = &t->nodes[c];
        gm = gm + ch->mass;
        cx = cx + ch->comx * ch->mass;
        cy = cy + ch->comy * ch->mass;
        next += 1 + ch->next;
    }
    if (gm > 0) { cx = cx / gm; cy = cy / gm; }
    nd.mass = gm; nd.comx = cx; nd.comy = cy; nd.next = next;
    t->nodes[idx] = nd;
    return 0;
}

static int64_t tree_build_internal(int32_t n, const float *x, const float *y, const float *m,
                                   const float bb[3], uint32_t leaf_cap, geo_node **out) {
    tree_builder t;
    t.leaf_cap = leaf_cap;
    t.cap = 4 * (int64_t)n + 16; /* Quadtree::root reserves 4N (quadtree.cpp:94) */
    t.count = 0;
    t.nodes = (geo_node *)malloc(sizeof(geo_node) * (size_t)t.cap);
    t.x = x; t.y = y; t.m = m;
    if (!t.nodes) return -1;
    int64_t r = tb_alloc(&t);
    if (r < 0 || tb_node(&t, r, bb[0], bb[1], bb[2], 0, (uint32_t)n) != 0) {
        free(t.nodes);
        return -1;
    }
    *out = t.nodes;
    return t.count;
}

int32_t geo_build_tree(int32_t n, const float *x, const float *y, const float *m,
                       const float bb[3], geo_node *nodes, int32_t cap) {
    return geo_build_tree_cap(n, x, y, m, bb, 0, nodes, cap);
}

int32_t geo_build_tree_cap(int32_t n, const float *x, const float *y, const float *m,
                           const float bb[3], uint32_t leaf_cap, geo_node *nodes, int32_t cap) {
    geo_node *tmp = NULL;
    int64_t c = tree_build_internal(n, x, y, m, bb, leaf_cap, &tmp);
    if (c < 0) return -1;
    int64_t k = c < cap ? c : cap;
    if (nodes && k > 0) memcpy(nodes, tmp, sizeof(geo_node) * (size_t)k);
    free(tmp);
    return (int32_t)c;
}

/* ------------------------------------------------------------- gravity --- */
/* Physics::calculateForceFromGrid, GE/src/Physics/physics.cpp:91-205
 * (temperature and periodic branches off: isTempEnabled=false,
 * infiniteDomain=true). */
void geo_bh_accel(int32_t n, const float *x, const float *y, const float *m,
                  const geo_node *nodes, int32_t nn, float G, float theta, float eps,
                  float *ax, float *ay, uint64_t *interactions) {
    const geo_domain none = {0.0f, 0.0f};
    geo_bh_accel_dom(n, x, y, m, nodes, nn, G, theta, eps, &none, ax, ay, interactions);
}

/* d -= domain * ((d > half) - (d < -half)), physics.cpp:117-120 */
static inline float wrap(float d, float size, float half) {
    return d - size * (float)((d > half) - (d < -half));
}

void geo_bh_accel_dom(int32_t n, const float *x, const float *y, const float *m,
                      const geo_node *nodes, int32_t nn, float G, float theta, float eps,
                      const geo_domain *dom, float *ax, float *ay, uint64_t *interactions) {
    const int per = dom->w > 0.0f;
    const float hw = dom->w * 0.5f, hh = dom->h * 0.5f;
    const float thetaSq = theta * theta;
    const float softSq = eps * eps;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int32_t i = 0; i < n; ++i) {
        float fx = 0.0f, fy = 0.0f;
        const float pmass = m[i];
        const float xi = x[i], yi = y[i];
        uint32_t idx = 0;
        uint64_t cnt = 0;
        while (idx < (uint32_t)nn) {
            const geo_node *g = &nodes[idx];
            float gmass = g->mass;
            if (gmass <= 0.0f) { idx += g->next + 1; continue; }
            float dx = g->comx - xi;
            float dy = g->comy - yi;
            if (per) { dx = wrap(dx, dom->w, hw); dy = wrap(dy, dom->h, hh); }
            float distSq = dx * dx + dy * dy + softSq;
            int leaf = g->next == 0; /* all subGrids empty */
            float sizeSq = g->size * g->size;
            if (sizeSq < thetaSq * distSq) {
                float inv = 1.0f / sqrtf(distSq);
                float inv2 = inv * inv;
                float inv3 = inv2 * inv;
                float fm = G * pmass * gmass * inv3;
                fx = fx + dx * fm;
                fy = fy + dy * fm;
                cnt++;
                idx += g->next + 1;
            } else if (leaf) {
                for (uint32_t j = g->start; j < g->end; ++j) {
                    if ((uint32_t)i == j) continue;
                    float px = x[j] - xi;
                    float py = y[j] - yi;
                    if (per) { px = wrap(px, dom->w, hw); py = wrap(py, dom->h, hh); }
                    float d2 = px * px + py * py + softSq;
                    float inv = 1.0f / sqrtf(d2);
                    float inv2 = inv * inv;
                    float inv3 = inv2 * inv;
                    float fm = G * pmass * m[j] * inv3;
                    fx = fx + px * fm;
                    fy = fy + py * fm;
                    cnt++;
                }
                idx += g->next + 1;
            } else {
                ++idx;
            }
        }
        ax[i] = fx / m[i];
        ay[i] = fy / m[i];
        total += cnt;
    }
    if (interactions) *interactions = total;
}

/* Physics::naiveGravity, GE/src/Physics/physics.cpp:392-434 */
void geo_direct_accel(int32_t n, const float *x, const float *y, const float *m,
                      float G, float eps, float *ax, float *ay) {
    const geo_domain none = {0.0f, 0.0f};
    geo_direct_accel_dom(n, x, y, m, G, eps, &none, ax, ay);
}

void geo_direct_accel_dom(int32_t n, const float *x, const float *y, const float *m,
                          float G, float eps, const geo_domain *dom, float *ax, float *ay) {
    const int per = dom->w > 0.0f;
    const float hw = dom->w * 0.5f, hh = dom->h * 0.5f;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 0; i < n; ++i) {
        float pix = x[i], piy = y[i];
        float tx = 0.0f, ty = 0.0f;
        for (int32_t j = 0; j < n; ++j) {
            float dx = x[j] - pix;
            float dy = y[j] - piy;
            if (per) { dx = wrap(dx, dom->w, hw); dy = wrap(dy, dom->h, hh); } /* physics.cpp:417-420 */
            float distSq = dx * dx + dy * dy + eps; /* softening not squared */
            float inv = 1.0f / sqrtf(distSq);
            float inv3 = inv * inv * inv;
            float factor = G * m[j] * inv3;
            tx = tx + dx * factor;
            ty = ty + dy * factor;
        }
        ax[i] = tx;
        ay[i] = ty;
    }
}

void geo_direct_fp64(int32_t n, const float *x, const float *y, const float *m,
                     double G, double eps, int32_t nt, const uint32_t *targets,
                     double *ax, double *ay) {
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t t = 0; t < nt; ++t) {
        uint32_t i = targets ? targets[t] : (uint32_t)t;
        double xi = x[i], yi = y[i];
        double sx = 0.0, sy = 0.0;
        for (int32_t j = 0; j < n; ++j) {
            if ((uint32_t)j == i) continue;
            double dx = (double)x[j] - xi;
            double dy = (double)y[j] - yi;
            double r2 = dx * dx + dy * dy + e2;
            double inv = 1.0 / sqrt(r2);
            double f = (double)m[j] * inv * inv * inv;
            sx += dx * f;
            sy += dy * f;
        }
        ax[t] = G * sx;
        ay[t] = G * sy;
    }
}

/* ---------------------------------------------------------- integrator --- */
/* Physics::integrateStart, GE/src/Physics/physics.cpp:1069-1112 */
void geo_integrate_start(geo_state *s, float dt, int32_t sph, float sph_max_vel) {
    const geo_domain none = {0.0f, 0.0f};
    geo_integrate_start_dom(s, dt, sph, sph_max_vel, &none);
}

void geo_integrate_start_dom(geo_state *s, float dt, int32_t sph, float sph_max_vel, const geo_domain *dom) {
    const float halfDt = dt * 0.5f;
    const float maxVSq = sph_max_vel * sph_max_vel;
    for (int32_t i = 0; i < s->n; ++i) {
        if (s->flags[i] & GEO_PINNED) continue;
        float pvx = s->vx[i], pvy = s->vy[i];
        float vx = s->vx[i] + s->ax[i] * halfDt;
        float vy = s->vy[i] + s->ay[i] * halfDt;
        if (sph) {
            float vSq = vx * vx + vy * vy;
            if (vSq > maxVSq) {
                float pSq = pvx * pvx + pvy * pvy;
                if (pSq > 0.00001f) {
                    float inv = sph_max_vel / sqrtf(pSq);
                    pvx = pvx * inv;
                    pvy = pvy * inv;
                }
                float inv = sph_max_vel / sqrtf(vSq);
                vx = vx * inv;
                vy = vy * inv;
            }
        }
        s->pvx[i] = pvx; s->pvy[i] = pvy;
        s->vx[i] = vx; s->vy[i] = vy;
        s->x[i] = s->x[i] + vx * dt;
        s->y[i] = s->y[i] + vy * dt;
        if (dom->w > 0.0f) { /* physics.cpp:1104-1110 */
            if (s->x[i] < 0.0f) s->x[i] += dom->w;
            else if (s->x[i] >= dom->w) s->x[i] -= dom->w;
            if (s->y[i] < 0.0f) s->y[i] += dom->h;
            else if (s->y[i] >= dom->h) s->y[i] -= dom->h;
        }
    }
}

/* Physics::integrateEnd, GE/src/Physics/physics.cpp:1114-1125 */
void geo_integrate_end(geo_state *s, float dt) {
    const float h = dt * 0.5f;
    for (int32_t i = 0; i < s->n; ++i) {
        if (s->flags[i] & GEO_PINNED) continue;
        s->vx[i] = s->vx[i] + s->ax[i] * h;
        s->vy[i] = s->vy[i] + s->ay[i] * h;
    }
}

/* ----------------------------------------------------------------- SPH --- */
typedef struct { int32_t cx, cy; uint32_t i; } cell_entry;

static int cell_cmp(const void *a, const void *b) {
    const cell_entry *p = (const cell_entry *)a, *q = (const cell_entry *)b;
    if (p->cx != q->cx) return p->cx < q->cx ? -1 : 1;
    if (p->cy != q->cy) return p->cy < q->cy ? -1 : 1;
    return p->i < q->i ? -1 : (p->i > q->i);
}

/* search semantics: mode 0 = NeighborSearchV2AVX2, mode 1 = NeighborSearchV2 */
typedef struct {
    int32_t mode;
    float cell, inv, r2;
    int32_t R;  /* cells searched on each side of the query cell */
} search_t;

static search_t make_search(int32_t mode, float cell, float radius) {
    search_t S;
    S.mode = mode;
    S.cell = cell;
    S.inv = 1.0f / cell;                  /* invCellSize (neighborSearch.h:443) */
    S.r2 = radius * radius;               /* searchRadius^2 (neighborSearch.h:553) */
    S.R = mode ? (int32_t)ceilf(radius * S.inv) : 1;  /* neighborSearch.h:552 */
    return S;
}

/* posToCellCoord: AVX2 floor(p / cell) (neighborSearch.h:631-635); V2
 * (int)(p * invCellSize), a truncation (neighborSearch.h:472-474) */
static inline int32_t cell_coord(float p, const search_t *S) {
    if (S->mode) return (int32_t)(p * S->inv);
    return (int32_t)floorf(p / S->cell);
}

typedef struct { cell_entry *e; int32_t n; const float *x, *y; search_t S; } cell_grid;

/* newGridAVX2 (neighborSearch.h:636-720) or newGrid (:481-541), restated as
 * a (cx, cy, index) ordered table; the candidate SET is the hashed table's. */
static cell_grid grid_build(int32_t n, const float *x, const float *y, const search_t *S) {
    cell_grid g;
    g.n = n;
    g.x = x;
    g.y = y;
    g.S = *S;
    g.e = (cell_entry *)malloc(sizeof(cell_entry) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; ++i) {
        g.e[i].cx = cell_coord(x[i], S);
        g.e[i].cy = cell_coord(y[i], S);
        g.e[i].i = (uint32_t)i;
    }
    qsort(g.e, (size_t)n, sizeof(cell_entry), cell_cmp);
    return g;
}

static int32_t grid_lower(const cell_grid *g, int32_t cx, int32_t cy) {
    int32_t lo = 0, hi = g->n;
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        const cell_entry *e = &g->e[mid];
        if (e->cx < cx || (e->cx == cx && e->cy < cy)) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* queryNeighborsAVX2 (neighborSearch.h:722-777): dx outer, dy inner, every
 * particle whose cell coordinates match; no distance filter; self included.
 * NeighborSearchV2::queryNeighbors (neighborSearch.h:548-578): R cells each
 * way and |p_j - q|^2 <= searchRadius^2 with dx = p_j.x - q.x. */
static int32_t grid_query(const cell_grid *g, float qx, float qy, uint32_t *buf, int32_t cap) {
    const search_t *S = &g->S;
    int32_t cx = cell_coord(qx, S), cy = cell_coord(qy, S);
    int32_t k = 0;
    for (int dx = -S->R; dx <= S->R; ++dx) {
        for (int dy = -S->R; dy <= S->R; ++dy) {
            int32_t nx = cx + dx, ny = cy + dy;
            int32_t p = grid_lower(g, nx, ny);
            while (p < g->n && g->e[p].cx == nx && g->e[p].cy == ny) {
                const uint32_t j = g->e[p].i;
                int take = 1;
                if (S->mode) {
                    float ddx = g->x[j] - qx, ddy = g->y[j] - qy;
                    take = ddx * ddx + ddy * ddy <= S->r2;
                }
                if (take) {
                    if (k < cap) buf[k] = j;
                    ++k;
                }
                ++p;
            }
        }
    }
    return k;
}

static int u32_cmp(const void *a, const void *b) {
    uint32_t p = *(const uint32_t *)a, q = *(const uint32_t *)b;
    return p < q ? -1 : (p > q);
}

int64_t geo_sph_candidates(int32_t n, const float *x, const float *y,
                           int32_t nq, const float *qx, const float *qy, float cell,
                           int64_t *off, uint32_t *idx, int64_t cap) {
    return geo_sph_candidates_mode(n, x, y, nq, qx, qy, cell, 0, 0.0f, off, idx, cap);
}

void geo_neighbor_amount(int32_t n, const float *x, const float *y, const uint32_t *flags, float cell,
                         float radius, uint32_t *counts) {
    const search_t S = make_search(1, cell, radius);
    cell_grid g = grid_build(n, x, y, &S);
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; ++i) {
        const uint32_t f = flags ? flags[i] : 0u;
        counts[i] = (f & (GEO_PINNED | GEO_DRAWN | GEO_DARK)) ? 0u : (uint32_t)grid_query(&g, x[i], y[i], buf, n);
    }
    free(buf);
    free(g.e);
}

int64_t geo_sph_candidates_mode(int32_t n, const float *x, const float *y,
                                int32_t nq, const float *qx, const float *qy, float cell,
                                int32_t mode, float radius, int64_t *off, uint32_t *idx, int64_t cap) {
    const search_t S = make_search(mode, cell, radius);
    cell_grid g = grid_build(n, x, y, &S);
    int64_t total = 0;
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t q = 0; q < nq; ++q) {
        off[q] = total;
        int32_t k = grid_query(&g, qx[q], qy[q], buf, n);
        qsort(buf, (size_t)k, sizeof(uint32_t), u32_cmp);
        if (idx && total + k <= cap) memcpy(idx + total, buf, sizeof(uint32_t) * (size_t)k);
        total += k;
    }
    off[nq] = total;
    free(buf);
    free(g.e);
    return total <= cap ? total : -1;
}

#ifndef GEO_CFLAGS
#define GEO_CFLAGS "unknown flags"
#endif
const char *geo_build_info(void) {
#if defined(__clang__)
    return "clang " __clang_version__ " " GEO_CFLAGS;
#elif defined(__GNUC__)
    return "gcc " __VERSION__ " " GEO_CFLAGS;
#else
    return "unknown compiler " GEO_CFLAGS;
#endif
}

#define PI_F 3.14159265358979323846f /* raylib PI (a float) */

/* SPH.h:38-56 in float arithmetic.  The calls are pow(float, float) with the
 * float raylib PI: the reference's precompiled header includes raymath.h, whose
 * <math.h> is the C++ wrapper that brings std::pow's float overload into the
 * global namespace, so pow(h, 4.0f) is powf(h, 4.0f) and every product and
 * quotient below is a float operation. */
void geo_sph_constants(float h, float out[3]) {
    out[0] = (PI_F * powf(h, 4.0f)) / 6.0f;  /* smoothingKernel volume */
    out[1] = -45.0f / (PI_F * powf(h, 6.0f)); /* spikyKernelDerivative scale */
    out[2] = 45.0f / (PI_F * powf(h, 6.0f));  /* smoothingKernelLaplacian */
}

/* SPH::spikyKernelDerivative, SPH.h:45-50: float scale times powf(h - r, 2) */
static inline float spiky_grad(float r, float h, float scale) {
    if (r >= h) return 0.0f;
    return scale * powf(h - r, 2.0f);
}

/* SPH::smoothingKernelCohesion, SPH.h:58-64 */
static inline float coh_kernel(float r, float h) {
    if (r >= h) return 0.0f;
    float q = r / h;
    return (1.0f - q) * (0.5f - q) * (0.5f - q) * 30.0f / (PI_F * h * h);
}

/* SPH::PCISPH with maxIter = 1 (SPH.cpp:152-281) and computeViscCohesionForces
 * (SPH.cpp:98-150); scatter form in index order. */
int32_t geo_sph_pcisph(geo_state *s, const geo_params *p) {
    const int32_t n = s->n;
    const float h = p->sph_h, h2 = h * h;
    float kc[3];
    geo_sph_constants(h, kc);
    const float volume = kc[0], spiky = kc[1], lap = kc[2];
    const float coeff = 0.5f * p->dt * p->dt;
    float *fx = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    float *fy = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    const search_t S = make_search(p->sph_search, p->sph_cell, p->sph_search_radius);
    cell_grid g = grid_build(n, s->x, s->y, &S);
    const uint32_t *fl = s->flags;
#define IS_SPH(k) ((fl[k] & GEO_SPH) != 0)
#define DRAWN(k) ((fl[k] & GEO_DRAWN) != 0)
#define PINNED(k) ((fl[k] & GEO_PINNED) != 0)

    /* viscosity + cohesion, SPH.cpp:98-150 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || PINNED(i) || DRAWN(i)) continue;
        int32_t k = grid_query(&g, s->x[i], s->y[i], buf, n);
        for (int32_t t = 0; t < k; ++t) {
            uint32_t j = buf[t];
            if (!IS_SPH(j) || DRAWN(j)) continue;
            if ((int32_t)j == i) continue;
            float dx = s->x[j] - s->x[i], dy = s->y[j] - s->y[i];
            float rSq = dx * dx + dy * dy;
            if (rSq >= h2) continue;
            float r = sqrtf(rSq < 1e-6f ? 1e-6f : rSq); /* std::max(rSq, 1e-6f) */
            float nrx = dx / r, nry = dy / r;
            float mJ = s->sph_mass[j] * p->sph_unit_mass;
            float lapW = lap; /* smoothingKernelLaplacian: constant for r < h */
            float dj = s->dens[j] < 0.001f ? 0.001f : s->dens[j]; /* std::max(dens, 0.001f) */
            float vfx = p->viscosity * s->visc[j] * mJ / dj * lapW * (s->vx[j] - s->vx[i]);
            float vfy = p->viscosity * s->visc[j] * mJ / dj * lapW * (s->vy[j] - s->vy[i]);
            float cohCoef = p->cohesion_coef * s->cohesion[i];
            float cf = coh_kernel(r, h);
            float cfx = cohCoef * mJ * cf * nrx;
            float cfy = cohCoef * mJ * cf * nry;
            fx[i] = fx[i] + (vfx + cfx);
            fy[i] = fy[i] + (vfy + cfy);
            fx[j] = fx[j] - (vfx + cfx);
            fy[j] = fy[j] - (vfy + cfy);
        }
    }
    for (int32_t i = 0; i < n; ++i) {
        s->press[i] = 0.0f;
        s->pfx[i] = 0.0f;
        s->pfy[i] = 0.0f;
    }
    /* predicted positions, SPH.cpp:173-181 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || DRAWN(i)) continue;
        float dxp = (fx[i] / s->sph_mass[i]) * coeff;
        float dyp = (fy[i] / s->sph_mass[i]) * coeff;
        s->ppx[i] = s->x[i] + dxp;
        s->ppy[i] = s->y[i] + dyp;
    }
    /* predicted density and pressure, SPH.cpp:183-219 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || DRAWN(i)) continue;
        float pd = 0.0f;
        int32_t k = grid_query(&g, s->ppx[i], s->ppy[i], buf, n);
        for (int32_t t = 0; t < k; ++t) {
            uint32_t j = buf[t];
            if (!IS_SPH(j) || DRAWN(j)) continue;
            float dx = s->ppx[i] - s->ppx[j], dy = s->ppy[i] - s->ppy[j];
            float rrSq = dx * dx + dy * dy;
            if (rrSq >= h * h) continue;
            float rr = sqrtf(rrSq);
            float mJ = s->sph_mass[j] * p->sph_unit_mass;
            float rho0 = 0.5f * (s->rest_dens[i] + s->rest_dens[j]);
            float w = (h - rr) * (h - rr) / volume;
            pd = pd + mJ * w / rho0;
        }
        s->pred_dens[i] = pd;
        float err = pd - s->rest_dens[i];
        float pt = p->delta * err;
        if (pt < 0.0f) pt = 0.0f;
        s->press_tmp[i] = pt;
        s->press[i] = s->press[i] + pt * s->stiff[i] * p->stiff_mult;
    }
    /* pressure force, SPH.cpp:221-264 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || DRAWN(i)) continue;
        int32_t k = grid_query(&g, s->ppx[i], s->ppy[i], buf, n);
        for (int32_t t = 0; t < k; ++t) {
            uint32_t j = buf[t];
            if ((int32_t)j == i) continue;
            if (!IS_SPH(j) || DRAWN(j)) continue;
            float dx = s->ppx[i] - s->ppx[j], dy = s->ppy[i] - s->ppy[j];
            float rr = sqrtf(dx * dx + dy * dy);
            if (rr < 1e-5f || rr >= h) continue;
            float gradW = spiky_grad(rr, h, spiky);
            float nx = dx / rr, ny = dy / rr;
            float avgP = 0.5f * (s->press[i] + s->press[j]);
            float avgD = 0.5f * (s->pred_dens[i] + s->pred_dens[j]);
            float mi = s->sph_mass[i] * p->sph_unit_mass;
            float mj = s->sph_mass[j] * p->sph_unit_mass;
            float mag = -(mi + mj) * avgP / (avgD < 0.01f ? 0.01f : avgD);
            float smax = s->sph_mass[i] < s->sph_mass[j] ? s->sph_mass[j] : s->sph_mass[i];
            float smin = s->sph_mass[j] < s->sph_mass[i] ? s->sph_mass[j] : s->sph_mass[i];
            float ratio = smax / smin;
            float sc = 8.0f / ratio;
            if (1.0f < sc) sc = 1.0f;
            mag = mag * sc;
            float pfx = mag * gradW * nx, pfy = mag * gradW * ny;
            fx[i] = fx[i] + pfx;
            fy[i] = fy[i] + pfy;
            fx[j] = fx[j] - pfx;
            fy[j] = fy[j] - pfy;
        }
    }
    /* pressF and acceleration, SPH.cpp:271-280 */
    for (int32_t i = 0; i < n; ++i) {
        s->pfx[i] = fx[i] / s->sph_mass[i];
        s->pfy[i] = fy[i] / s->sph_mass[i];
        if (!PINNED(i)) {
            s->ax[i] = s->ax[i] + s->pfx[i];
            s->ay[i] = s->ay[i] + s->pfy[i];
        }
    }
#undef IS_SPH
#undef DRAWN
#undef PINNED
    free(g.e);
    free(buf);
    free(fx);
    free(fy);
    return 0;
}

void geo_sph_ground(geo_state *s, float radius, float w, float h, float vertical_gravity, float friction) {
    const float damp = -0.1f; /* SPH::boundDamping, SPH.h:18 */
    for (int32_t i = 0; i < s->n; ++i) {
        if (s->flags[i] & GEO_PINNED) continue;
        s->ay[i] += vertical_gravity;
        if (s->x[i] - radius < 0.0f) { /* left */
            s->vx[i] *= damp; s->vy[i] *= 1.0f - friction; s->x[i] = radius;
        }
        if (s->x[i] + radius > w) { /* right */
            s->vx[i] *= damp; s->vy[i] *= 1.0f - friction; s->x[i] = w - radius;
        }
        if (s->y[i] - radius < 0.0f) { /* bottom */
            s->vy[i] *= damp; s->vx[i] *= 1.0f - friction; s->y[i] = radius;
        }
        if (s->y[i] + radius > h) { /* top */
            s->vy[i] *= damp; s->vx[i] *= 1.0f - friction; s->y[i] = h - radius;
        }
    }
}

/* ---------------------------------------------------------- full step --- */
static void permute_f(float *a, const uint32_t *perm, int32_t n, float *tmp) {
    if (!a) return;
    for (int32_t i = 0; i < n; ++i) tmp[i] = a[perm[i]];
    memcpy(a, tmp, sizeof(float) * (size_t)n);
}
static void permute_u(uint32_t *a, const uint32_t *perm, int32_t n, uint32_t *tmp) {
    if (!a) return;
    for (int32_t i = 0; i < n; ++i) tmp[i] = a[perm[i]];
    memcpy(a, tmp, sizeof(uint32_t) * (size_t)n);
}

/* Morton::sortParticlesByMortonKey (morton.cpp:40-63): gather every field */
static void permute_state(geo_state *s, const uint32_t *perm) {
    int32_t n = s->n;
    float *t = (float *)malloc(sizeof(float) * (size_t)n);
    float *fs[] = {s->x, s->y, s->vx, s->vy, s->ax, s->ay, s->pvx, s->pvy, s->m,
                   s->sph_mass, s->rest_dens, s->stiff, s->visc, s->cohesion, s->dens,
                   s->press, s->press_tmp, s->pred_dens, s->ppx, s->ppy, s->pfx, s->pfy,
                   s->temp, s->ke, s->prev_ke};
    for (size_t k = 0; k < sizeof(fs) / sizeof(fs[0]); ++k) permute_f(fs[k], perm, n, t);
    permute_u(s->flags, perm, n, (uint32_t *)t);
    permute_u(s->id, perm, n, (uint32_t *)t);
    permute_u(s->label, perm, n, (uint32_t *)t);
    free(t);
}

/* updateScene() 2D physics ordering, GE/src/globalLogic.cpp:1155-1372 */
int32_t geo_step(geo_state *s, const geo_params *p, int32_t nsteps) {
    const int32_t n = s->n;
    if (n <= 0) return 0;
    geo_domain dom = {0.0f, 0.0f};
    if (p->periodic) { dom.w = p->domain_w; dom.h = p->domain_h; }
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint32_t *perm = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int32_t step = 0; step < nsteps; ++step) {
        float dt0 = step == 0 ? p->dt_start : p->dt;
        geo_node *nodes = NULL;
        int64_t nn = 0;
        if (dt0 != 0.0f) {
            geo_integrate_start_dom(s, dt0, p->sph, p->sph_max_vel, &dom);
            float bb[3];
            geo_bbox(n, s->x, s->y, bb);
            geo_morton_keys(n, s->x, s->y, bb, keys);
            geo_sort_perm(n, keys, perm);
            permute_state(s, perm);
            if (!p->naive) {
                nn = tree_build_internal(n, s->x, s->y, s->m, bb, 0, &nodes);
                if (nn < 0) { free(keys); free(perm); return -2; }
            }
        }
        if (p->dt != 0.0f) {
            for (int32_t i = 0; i < n; ++i) { s->ax[i] = 0.0f; s->ay[i] = 0.0f; }
            const int temp_on = p->temperature && s->temp;
            /* the walk runs for gravity or for the heat transfer
             * (globalLogic.cpp:1316); the direct sum moves no heat */
            if (p->G != 0.0f || temp_on) {
                if (p->naive || n < p->direct_below || !nodes) {
                    geo_direct_accel_dom(n, s->x, s->y, s->m, p->G, p->softening, &dom, s->ax, s->ay);
                } else {
                    geo_bh_accel_dom(n, s->x, s->y, s->m, nodes, (int32_t)nn, p->G, p->theta,
                                     p->softening, &dom, s->ax, s->ay, NULL);
                    if (temp_on) {
                        float *gt = (float *)malloc(sizeof(float) * (size_t)nn);
                        geo_node_temps(nodes, (int32_t)nn, s->temp, gt);
                        geo_walk_heat(n, s->x, s->y, nodes, (int32_t)nn, gt, p->theta, p->softening,
                                      p->heat_conductivity, p->dt, &dom, s->temp);
                        free(gt);
                    }
                }
            }
            if (p->sph) {
                geo_sph_pcisph(s, p);
                if (p->sph_ground)
                    geo_sph_ground(s, p->sph_h, p->domain_w, p->domain_h, p->vertical_gravity,
                                   p->boundary_friction);
            }
            if (temp_on)
                geo_temperature(s, p->materials, p->nmaterials, p->ambient_temp, p->ambient_heat_rate, p->dt);
            geo_integrate_end(s, p->dt);
        }
        free(nodes);
    }
    free(keys);
    free(perm);
    return 0;
}

/* --------------------------------------------------------- temperature --- */
/* Node::computeLeafMass / computeInternalMass's gridTemp
 * (GE/include/Physics/quadtree.h:53-100).  Children are allocated after their
 * parent, so a reverse sweep meets every child first. */
void geo_node_temps(const geo_node *nodes, int32_t nn, const float *temp, float *gtemp) {
    static const int order[4] = {0, 2, 1, 3};
    for (int32_t k = nn - 1; k >= 0; --k) {
        const geo_node *g = &nodes[k];
        float t = 0.0f;
        if (g->child[0] == U32MAX && g->child[1] == U32MAX && g->child[2] == U32MAX && g->child[3] == U32MAX) {
            for (uint32_t i = g->start; i < g->end; ++i) t = t + temp[i];
        } else {
            for (int q = 0; q < 4; ++q) {
                const uint32_t c = g->child[order[q]];
                if (c != U32MAX) t = t + gtemp[c];
            }
        }
        gtemp[k] = t;
    }
}

/* the isTempEnabled branch of calculateForceFromGridAVX2
 * (GE/src/Physics/physics.cpp:261-276) over the same traversal */
void geo_walk_heat(int32_t n, const float *x, const float *y, const geo_node *nodes, int32_t nn,
                   const float *gtemp, float theta, float eps, float conductivity, float dt,
                   const geo_domain *dom, float *temp) {
    const int per = dom->w > 0.0f;
    const float hw = dom->w * 0.5f, hh = dom->h * 0.5f;
    const float thetaSq = theta * theta;
    const float softSq = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 0; i < n; ++i) {
        float ti = temp[i];
        const float xi = x[i], yi = y[i];
        uint32_t idx = 0;
        while (idx < (uint32_t)nn) {
            const geo_node *g = &nodes[idx];
            if (g->mass <= 0.0f) { idx += g->next + 1; continue; }
            float dx = g->comx - xi;
            float dy = g->comy - yi;
            if (per) { dx = wrap(dx, dom->w, hw); dy = wrap(dy, dom->h, hh); }
            const float distSq = dx * dx + dy * dy + softSq;
            const int leaf = g->next == 0;
            const float sizeSq = g->size * g->size;
            if (sizeSq < thetaSq * distSq) {
                const uint32_t count = g->end - g->start;
                if (count > 0) {
                    const float avg = gtemp[idx] / (float)count;
                    const float diff = avg - ti;
                    const float inv = 1.0f / sqrtf(distSq);
                    float dist = 0.0f;
                    if (distSq > 1e-16f) dist = 1.0f / inv;
                    if (dist > 1e-8f) {
                        const float heat = conductivity * diff / dist;
                        ti = ti + heat * dt;
                    }
                }
                idx += g->next + 1;
            } else if (leaf) {
                idx += g->next + 1;
            } else {
                ++idx;
            }
        }
        temp[i] = ti;
    }
}

static const geo_material *find_material(const geo_material *mats, int32_t nmat, uint32_t id) {
    for (int32_t k = 0; k < nmat; ++k)
        if (mats[k].id == id) return &mats[k];
    return NULL;
}

/* Physics::createConstraints' automatic branch (physics.cpp:654-766).  The
 * constraintMap check never fires here: the map is cleared for every
 * solidified particle (physics.cpp:692) and each candidate appears once in
 * its query, so no pair is skipped as already present. */
int32_t geo_create_constraints(geo_state *s, const geo_material *mats, int32_t nmat, int32_t mode, float cell,
                               float radius, geo_constraint *out, int32_t cap) {
    const int32_t n = s->n;
    const search_t S = make_search(mode, cell, radius);
    int32_t *cx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *cy = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t j = 0; j < n; ++j) {
        cx[j] = cell_coord(s->x[j], &S);
        cy[j] = cell_coord(s->y[j], &S);
    }
    int32_t m = 0, over = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (!(s->flags[i] & GEO_SOLIDIFIED)) continue;
        s->flags[i] &= ~GEO_SOLIDIFIED; /* pi.hasSolidified = false (physics.cpp:693) */
        const geo_material *mi = s->label ? find_material(mats, nmat, s->label[i]) : NULL;
        const int32_t qx = cell_coord(s->x[i], &S), qy = cell_coord(s->y[i], &S);
        for (int32_t dx = -S.R; dx <= S.R; ++dx) {
            for (int32_t dy = -S.R; dy <= S.R; ++dy) {
                for (int32_t j = 0; j < n; ++j) {
                    if (cx[j] != qx + dx || cy[j] != qy + dy) continue;
                    if (S.mode) { /* NeighborSearchV2's radius filter on the candidate */
                        const float fx = s->x[j] - s->x[i], fy = s->y[j] - s->y[i];
                        if (!(fx * fx + fy * fy <= S.r2)) continue;
                    }
                    if (j == i) continue;
                    const float ex = s->x[j] - s->x[i], ey = s->y[j] - s->y[i];
                    const float dist_sq = ex * ex + ey * ey; /* glm::dot */
                    if (dist_sq > 12.0f) continue;
                    const geo_material *mj = s->label ? find_material(mats, nmat, s->label[j]) : NULL;
                    if (mi && mj && mi->cold_point == 0.0f && mj->cold_point != 0.0f) continue;
                    if (!(s->id[i] < s->id[j])) continue;
                    geo_constraint k;
                    memset(&k, 0, sizeof k);
                    k.id1 = s->id[i];
                    k.id2 = s->id[j];
                    k.rest_length = k.original_length = sqrtf(dist_sq); /* glm::distance */
                    k.stiffness = (mi && mj) ? mi->constraint_stiffness : 60.0f;
                    k.resistance = (mi && mj) ? (mi->constraint_resistance + mj->constraint_resistance) * 0.5f : 0.6f;
                    k.plasticity_point =
                        (mi && mj) ? (mi->constraint_plastic_point + mj->constraint_plastic_point) * 0.5f : 0.6f;
                    if (m < cap) out[m] = k;
                    else over = 1;
                    ++m;
                }
            }
        }
    }
    free(cx);
    free(cy);
    return over ? -1 : m;
}

/* Physics::temperatureCalculation (GE/src/Physics/physics.cpp:567-652);
 * UpdateVariables::particleBaseMass = 8.5e9 (GE/src/parameters.cpp:44) */
void geo_temperature(geo_state *s, const geo_material *mats, int32_t nmat, float ambient, float rate, float dt) {
    const float base_mass = 8500000000.0f;
    for (int32_t i = 0; i < s->n; ++i) {
        const geo_material *mt = find_material(mats, nmat, s->label ? s->label[i] : 0u);
        const float K = mt ? mt->heat_conductivity : 0.05f;
        const float v = sqrtf(s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i]);
        const float pv = sqrtf(s->pvx[i] * s->pvx[i] + s->pvy[i] * s->pvy[i]);
        const float sm = s->sph_mass[i];
        const float ke = 0.5f * sm * v * v;
        const float pke = 0.5f * sm * pv * pv;
        s->ke[i] = ke;
        s->prev_ke[i] = pke;
        const float q = fabsf(ke - pke);
        const float dT = q / (2.0f * K * sm + 1.0f);
        float t = s->temp[i] + dT;
        const float diff = t - ambient;
        const float cool = -(K * rate) * diff * dt;
        t = t + cool;
        s->temp[i] = t;
        if (!mt) continue;
        uint32_t f = s->flags[i];
        int hot = (f & GEO_HOT_POINT) != 0, solid = (f & GEO_SOLIDIFIED) != 0;
        if (t >= mt->hot_point) {
            s->sph_mass[i] = mt->hot_mass_mult;
            s->rest_dens[i] = mt->hot_rest_dens;
            s->stiff[i] = mt->hot_stiff;
            s->visc[i] = mt->hot_visc;
            s->cohesion[i] = mt->hot_cohesion;
            if (mt->cold_point == 0.0f) hot = 1;
        } else if (t <= mt->cold_point) {
            s->sph_mass[i] = mt->cold_mass_mult;
            s->rest_dens[i] = mt->cold_rest_dens;
            s->stiff[i] = mt->cold_stiff;
            s->visc[i] = mt->cold_visc;
            s->cohesion[i] = mt->cold_cohesion;
        } else {
            s->sph_mass[i] = mt->mass_mult;
            s->rest_dens[i] = mt->rest_dens;
            s->stiff[i] = mt->stiff;
            s->visc[i] = mt->visc;
            s->cohesion[i] = mt->cohesion;
            if (mt->cold_point != 0.0f) hot = 1;
        }
        s->m[i] = base_mass * s->sph_mass[i];
        if (mt->cold_point == 0.0f) {
            if (t <= mt->hot_point && hot) { solid = 1; hot = 0; }
        } else {
            if (t <= mt->cold_point && hot) { solid = 1; hot = 0; }
        }
        f &= ~(GEO_HOT_POINT | GEO_SOLIDIFIED);
        s->flags[i] = f | (hot ? GEO_HOT_POINT : 0u) | (solid ? GEO_SOLIDIFIED : 0u);
    }
}

/* --------------------------------------------------------- constraints --- */
static int64_t index_of_id(const geo_state *s, uint32_t id) {
    for (int32_t i = 0; i < s->n; ++i)
        if (s->id[i] == id) return i;
    return -1;
}

typedef struct { uint32_t color, index; } order_item;
static int order_cmp(const void *a, const void *b) {
    const order_item *x = (const order_item *)a, *y = (const order_item *)b;
    if (x->color != y->color) return (x->color > y->color) - (x->color < y->color);
    return (x->index > y->index) - (x->index < y->index);
}

int32_t geo_constraints(geo_state *s, geo_constraint *cons, uint32_t *color, int32_t m, const geo_cparams *p,
                        const geo_material *mats, int32_t nmat, uint32_t *frame_count) {
    /* idToIndexTable and the removal of constraints with a missing end or
     * already broken (physics.cpp:777-817) */
    int64_t *ia = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    int64_t *ib = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    int32_t k = 0;
    for (int32_t c = 0; c < m; ++c) {
        const int64_t a = index_of_id(s, cons[c].id1), b = index_of_id(s, cons[c].id2);
        if (a < 0 || b < 0 || cons[c].is_broken) continue;
        cons[k] = cons[c];
        color[k] = color[c];
        ia[k] = a;
        ib[k] = b;
        ++k;
    }
    m = k;
    if (m == 0) { free(ia); free(ib); return 0; }
    *frame_count += 1;
    /* the triangle rule (physics.cpp:823-862): adjacency of unbroken constraints */
    if (*frame_count % 5 == 0) {
        uint32_t **adj = (uint32_t **)calloc((size_t)s->n, sizeof(uint32_t *));
        uint32_t *deg = (uint32_t *)calloc((size_t)s->n, sizeof(uint32_t));
        uint32_t *cap = (uint32_t *)calloc((size_t)s->n, sizeof(uint32_t));
        for (int32_t c = 0; c < m; ++c) {
            if (cons[c].is_broken) continue;
            const int64_t e[2] = {ia[c], ib[c]};
            for (int q = 0; q < 2; ++q) {
                const int64_t v = e[q], o = e[1 - q];
                if (deg[v] == cap[v]) {
                    cap[v] = cap[v] ? 2 * cap[v] : 4;
                    adj[v] = (uint32_t *)realloc(adj[v], sizeof(uint32_t) * cap[v]);
                }
                adj[v][deg[v]++] = (uint32_t)o;
            }
        }
        for (int32_t c = 0; c < m; ++c) {
            if (cons[c].is_broken) continue;
            const int64_t A = ia[c], B = ib[c];
            int tri = 0;
            for (uint32_t x = 0; x < deg[A] && !tri; ++x)
                for (uint32_t y = 0; y < deg[B]; ++y)
                    if (adj[A][x] == adj[B][y]) { tri = 1; break; }
            if (!tri) cons[c].is_broken = 1;
        }
        for (int32_t v = 0; v < s->n; ++v) free(adj[v]);
        free(adj); free(deg); free(cap);
    }
    order_item *ord = (order_item *)malloc(sizeof(order_item) * (size_t)m);
    for (int32_t c = 0; c < m; ++c) { ord[c].color = color[c]; ord[c].index = (uint32_t)c; }
    qsort(ord, (size_t)m, sizeof(order_item), order_cmp);
    for (int step = 0; step < 15; ++step) {
        for (int32_t t = 0; t < m; ++t) {
            const int32_t c = (int32_t)ord[t].index;
            geo_constraint *K = &cons[c];
            if (K->is_broken) continue;
            const int64_t i = ia[c], j = ib[c];
            const geo_material *mi = s->label ? find_material(mats, nmat, s->label[i]) : NULL;
            const geo_material *mj = s->label ? find_material(mats, nmat, s->label[j]) : NULL;
            float dx = s->x[j] - s->x[i], dy = s->y[j] - s->y[i];
            if (p->periodic) {
                dx = fmodf(dx + p->domain_w * 1.5f, p->domain_w) - p->domain_w * 0.5f;
                dy = fmodf(dy + p->domain_h * 1.5f, p->domain_h) - p->domain_h * 0.5f;
            }
            const float len = sqrtf(dx * dx + dy * dy);  /* glm::length */
            if (len < 0.0001f) continue;
            const float dirx = dx / len, diry = dy / len;
            K->displacement = len - K->rest_length;
            if (!p->unbreakable && mi && mj) {
                if (!mi->is_plastic || !mj->is_plastic) {
                    K->is_plastic = 0;
                    const float limit = (K->resistance * p->resistance_mult) * K->rest_length;
                    if (fabsf(K->displacement) >= limit) K->is_broken = 1;
                } else {
                    K->is_plastic = 1;
                    if (fabsf(K->displacement) >= K->plasticity_point * K->original_length)
                        K->rest_length += K->displacement;
                    const float brk = (K->original_length + K->original_length * (K->resistance * p->resistance_mult)) *
                                      (mi->constraint_plastic_point_mult + mj->constraint_plastic_point_mult) * 0.5f;
                    if (fabsf(K->rest_length) >= brk) K->is_broken = 1;
                }
                if ((s->flags[i] | s->flags[j]) & GEO_HOT_POINT) K->is_broken = 1;
            }
            if (p->forces && !K->is_broken) {
                const float m1 = s->m[i], m2 = s->m[j];
                const float sd = K->stiffness * K->displacement;
                const float sfx = ((sd * dirx) * m1) * p->stiffness_mult, sfy = ((sd * diry) * m1) * p->stiffness_mult;
                const float rvx = s->vx[j] - s->vx[i], rvy = s->vy[j] - s->vy[i];
                const float dd = -p->damping * (rvx * dirx + rvy * diry);
                const float dfx = (dd * dirx) * m1, dfy = (dd * diry) * m1;
                const float tx = sfx + dfx, ty = sfy + dfy;
                s->ax[i] = s->ax[i] + tx / m1;
                s->ay[i] = s->ay[i] + ty / m1;
                s->ax[j] = s->ax[j] - tx / m2;
                s->ay[j] = s->ay[j] - ty / m2;
                const float cf = K->stiffness * p->correction_ratio * p->stiffness_mult;
                const float cx = (dirx * K->displacement) * cf, cy = (diry * K->displacement) * cf;
                const float ms = m1 + m2;
                const float wi = m2 / ms, wj = m1 / ms;
                s->x[i] = s->x[i] + cx * wi;
                s->y[i] = s->y[i] + cy * wi;
                s->x[j] = s->x[j] - cx * wj;
                s->y[j] = s->y[j] - cy * wj;
            }
        }
    }
    free(ord);
    free(ia);
    free(ib);
    return m;
}

/* -------------------------------------------------------------- merger --- */
static void swap_f(float *a, int64_t i, int64_t j) { if (a) { float t = a[i]; a[i] = a[j]; a[j] = t; } }
static void swap_u(uint32_t *a, int64_t i, int64_t j) { if (a) { uint32_t t = a[i]; a[i] = a[j]; a[j] = t; } }
static void swap_i(int32_t *a, int64_t i, int64_t j) { if (a) { int32_t t = a[i]; a[i] = a[j]; a[j] = t; } }

static void swap_particles(geo_state *s, float *size, float *prev_size, int32_t *nb, int64_t i, int64_t j) {
    float *fs[] = {s->x, s->y, s->vx, s->vy, s->ax, s->ay, s->pvx, s->pvy, s->m,
                   s->sph_mass, s->rest_dens, s->stiff, s->visc, s->cohesion, s->dens,
                   s->press, s->press_tmp, s->pred_dens, s->ppx, s->ppy, s->pfx, s->pfy,
                   s->temp, s->ke, s->prev_ke, size, prev_size};
    for (size_t k = 0; k < sizeof(fs) / sizeof(fs[0]); ++k) swap_f(fs[k], i, j);
    swap_u(s->flags, i, j);
    swap_u(s->id, i, j);
    swap_u(s->label, i, j);
    swap_i(nb, i, j);
}

static int cmp_i64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

int32_t geo_merge(geo_state *s, float *size, float *prev_size, int32_t *neighbors, const geo_mparams *p) {
    const int32_t n = s->n;
    if (n == 0) return 0;
    const float cell = p->cell_size;
    float *rad = (float *)malloc(sizeof(float) * (size_t)n);
    for (int32_t i = 0; i < n; ++i) rad[i] = size[i] * p->texture_half_size * p->size_multiplier;
    /* UpdateNeighbors */
    int32_t *act = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t na = 0;
    for (int32_t i = 0; i < n; ++i)
        if (!(s->flags[i] & GEO_DARK)) act[na++] = i;
    uint32_t *nb_off = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    uint32_t *list = NULL;
    if (na > 0) {
        float minX = FLT_MAX, minY = FLT_MAX, maxX = -FLT_MAX, maxY = -FLT_MAX;
        for (int32_t k = 0; k < na; ++k) {
            const float x = s->x[act[k]], y = s->y[act[k]];
            if (x < minX) minX = x;
            if (y < minY) minY = y;
            if (x > maxX) maxX = x;
            if (y > maxY) maxY = y;
        }
        minX -= cell; minY -= cell; maxX += cell; maxY += cell;
        int gw = (int)((maxX - minX) / cell) + 1, gh = (int)((maxY - minY) / cell) + 1;
        if (gw < 1) gw = 1;
        if (gh < 1) gh = 1;
        const int64_t ncell = (int64_t)gw * gh;
        int32_t *cx = (int32_t *)malloc(sizeof(int32_t) * (size_t)na), *cy = (int32_t *)malloc(sizeof(int32_t) * (size_t)na);
        int64_t *cstart = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
        for (int32_t k = 0; k < na; ++k) {
            int a = (int)((s->x[act[k]] - minX) / cell), b = (int)((s->y[act[k]] - minY) / cell);
            a = a < 0 ? 0 : (a > gw - 1 ? gw - 1 : a);
            b = b < 0 ? 0 : (b > gh - 1 ? gh - 1 : b);
            cx[k] = a; cy[k] = b;
            cstart[(int64_t)b * gw + a + 1] += 1;
        }
        for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
        int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
        memcpy(cur, cstart, sizeof(int64_t) * (size_t)ncell);
        int32_t *cpart = (int32_t *)malloc(sizeof(int32_t) * (size_t)na);
        for (int32_t k = 0; k < na; ++k) cpart[cur[(int64_t)cy[k] * gw + cx[k]]++] = act[k];
        const float dr2 = p->density_radius * p->density_radius;
        /* two passes: counts (neighbors), then the lists of ids in scan order */
        for (int pass = 0; pass < 2; ++pass) {
            for (int32_t k = 0; k < na; ++k) {
                const int32_t i = act[k];
                uint32_t cnt = 0;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int a = cx[k] + dx, b = cy[k] + dy;
                        if (a < 0 || a >= gw || b < 0 || b >= gh) continue;
                        const int64_t c = (int64_t)b * gw + a;
                        for (int64_t q = cstart[c]; q < cstart[c + 1]; ++q) {
                            const int32_t j = cpart[q];
                            if (j == i) continue;
                            const float ex = s->x[i] - s->x[j], ey = s->y[i] - s->y[j];
                            if (ex * ex + ey * ey < dr2) {
                                if (pass == 1) list[nb_off[i] + cnt] = (uint32_t)j;  /* indices stand in for ids */
                                ++cnt;
                            }
                        }
                    }
                if (pass == 0) neighbors[i] = (int32_t)cnt;
            }
            if (pass == 0) {
                for (int32_t i = 0; i < n; ++i) nb_off[i + 1] = nb_off[i] + (uint32_t)(neighbors[i] > 0 ? neighbors[i] : 0);
                list = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nb_off[n] + 1));
            }
        }
        free(cx); free(cy); free(cstart); free(cur); free(cpart);
    }
    /* mergerSolver */
    uint8_t *del = (uint8_t *)calloc((size_t)n, 1);
    int64_t *dl = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t nd = 0;
    for (int32_t i = n - 1; i >= 0 && na > 0; --i) {
        if ((s->flags[i] & GEO_DARK) || del[i]) continue;
        for (int32_t t = neighbors[i] - 1; t >= 0; --t) {
            const int32_t j = (int32_t)list[nb_off[i] + (uint32_t)t];
            if (j == i) continue;
            if ((s->flags[j] & GEO_DARK) || del[j]) continue;
            const float dx = s->x[j] - s->x[i], dy = s->y[j] - s->y[i];
            const float d2 = dx * dx + dy * dy;
            const float cr = rad[i] + rad[j];
            if (d2 <= cr * cr) {
                const float mP = s->m[i], mN = s->m[j];
                const float a1 = prev_size[i] * prev_size[i], a2 = prev_size[j] * prev_size[j];
                const float full = sqrtf(a1 + a2);
                const float mx = prev_size[i] < prev_size[j] ? prev_size[j] : prev_size[i];
                if (mP >= mN) {
                    s->m[i] = mP + mN;
                    s->vx[i] = (s->vx[i] * mP + s->vx[j] * mN) / s->m[i];
                    s->vy[i] = (s->vy[i] * mP + s->vy[j] * mN) / s->m[i];
                    prev_size[i] = mx + (full - mx) * 0.25f;
                    del[j] = 1;
                    dl[nd++] = j;
                } else {
                    s->m[j] = mP + mN;
                    s->vx[j] = (s->vx[j] * mN + s->vx[i] * mP) / s->m[j];
                    s->vy[j] = (s->vy[j] * mN + s->vy[i] * mP) / s->m[j];
                    prev_size[j] = mx + (full - mx) * 0.25f;
                    del[i] = 1;
                    dl[nd++] = i;
                }
                break;
            }
        }
    }
    /* sort, unique, then swap with the back and pop, last index first */
    qsort(dl, (size_t)nd, sizeof(int64_t), cmp_i64);
    int64_t u = 0;
    for (int64_t k = 0; k < nd; ++k)
        if (u == 0 || dl[k] != dl[u - 1]) dl[u++] = dl[k];
    int64_t size_now = n;
    for (int64_t k = u - 1; k >= 0; --k) {
        const int64_t idx = dl[k];
        if (idx >= size_now) continue;
        swap_particles(s, size, prev_size, neighbors, idx, size_now - 1);
        --size_now;
    }
    s->n = (int32_t)size_now;
    free(rad); free(act); free(nb_off); free(list); free(del); free(dl);
    return (int32_t)size_now;
}
