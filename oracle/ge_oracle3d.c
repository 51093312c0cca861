/*
 * ge_oracle3d.c -- CPU restatement of Galaxy Engine's 3D physics step (gravity and SPH3D).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP 3D path); the product
 * library never links it.  Built into libge_oracle.so with the flags of
 * oracle/Makefile (-ffp-contract=off).  GE/ = GalaxyEngine/.
 */
#include "ge_oracle3d.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ge_oracle.h"

#define U32MAX 0xFFFFFFFFu

static inline float fmin_glm(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_glm(float a, float b) { return (a < b) ? b : a; }

/* boundingBox3D, GE/src/globalLogic.cpp:1465-1477 */
void geo3_bbox(int32_t n, const float *x, const float *y, const float *z, float out[4]) {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int32_t i = 0; i < n; ++i) {
        mn[0] = fmin_glm(mn[0], x[i]); mx[0] = fmax_glm(mx[0], x[i]);
        mn[1] = fmin_glm(mn[1], y[i]); mx[1] = fmax_glm(mx[1], y[i]);
        mn[2] = fmin_glm(mn[2], z[i]); mx[2] = fmax_glm(mx[2], z[i]);
    }
    float size = fmax_glm(fmax_glm(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    float half = size * 0.5f;
    for (int k = 0; k < 3; ++k) out[k] = (mn[k] + mx[k]) * 0.5f - half;
    out[3] = size;
}

/* Morton::scaleToGrid3D, morton.cpp:67-75 */
uint64_t geo3_scale_to_grid(float pos, float lo, float hi) {
    if (hi <= lo) return 0;
    float c = (pos < lo) ? lo : ((hi < pos) ? hi : pos);
    float norm = (c - lo) / (hi - lo);
    uint64_t s = (uint64_t)(norm * 2097151.0f);
    return s < 2097151u ? s : 2097151u;
}

/* Morton::spreadBits3D, morton.cpp:77-86: bit i of v to bit 3i */
uint64_t geo3_spread_bits(uint64_t v) {
    uint64_t r = 0;
    for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ull) << (3 * i);
    return r;
}

/* Morton::morton3D, morton.cpp:88-90 */
uint64_t geo3_morton(uint64_t ix, uint64_t iy, uint64_t iz) {
    return geo3_spread_bits(ix) | (geo3_spread_bits(iy) << 1) | (geo3_spread_bits(iz) << 2);
}

/* Morton::computeMortonKeys3D, morton.cpp:94-113 (no minimum extent) */
void geo3_morton_keys(int32_t n, const float *x, const float *y, const float *z, const float bb[4],
                      uint64_t *keys) {
    const float mxx = bb[0] + bb[3], mxy = bb[1] + bb[3], mxz = bb[2] + bb[3];
    for (int32_t i = 0; i < n; ++i)
        keys[i] = geo3_morton(geo3_scale_to_grid(x[i], bb[0], mxx), geo3_scale_to_grid(y[i], bb[1], mxy),
                              geo3_scale_to_grid(z[i], bb[2], mxz));
}

/* ------------------------------------------------------------- octree --- */
typedef struct {
    geo3_node *nodes;
    int64_t count, cap;
    const float *x, *y, *z, *m;
    uint32_t leaf_cap;
} tree3;

/* std::partition_point's halving search with predicate v[i] < mid */
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

static int64_t t3_alloc(tree3 *t) {
    if (t->count == t->cap) {
        int64_t nc = t->cap * 2 + 16;
        geo3_node *nn = (geo3_node *)realloc(t->nodes, sizeof(geo3_node) * (size_t)nc);
        if (!nn) return -1;
        t->nodes = nn;
        t->cap = nc;
    }
    return t->count++;
}

/* Node3D::Node3D (quadtree.cpp:109-129), subGridMaker3D (:146-226),
 * computeLeafMass3D (quadtree.h:159-174), computeInternalMass3D
 * (quadtree.cpp:228-253), calculateNextNeighbor3D (:255-273) */
static int t3_node(tree3 *t, int64_t idx, float px, float py, float pz, float size, uint32_t s, uint32_t e) {
    geo3_node nd;
    nd.posx = px; nd.posy = py; nd.posz = pz; nd.size = size;
    nd.start = s; nd.end = e; nd.next = 0;
    nd.mass = 0.0f; nd.comx = nd.comy = nd.comz = 0.0f;
    for (int q = 0; q < 8; ++q) nd.child[q] = U32MAX;
    const uint32_t cnt = e - s;
    const int leaf = t->leaf_cap ? (cnt <= t->leaf_cap || size <= 0.000001f)
                                 : (((cnt <= 4 && size <= 2.0f) || cnt == 1) || size <= 0.000001f);
    if (leaf) {
        float gm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        for (uint32_t i = s; i < e; ++i) {
            gm = gm + t->m[i];
            cx = cx + t->x[i] * t->m[i];
            cy = cy + t->y[i] * t->m[i];
            cz = cz + t->z[i] * t->m[i];
        }
        if (gm > 0) { cx = cx / gm; cy = cy / gm; cz = cz / gm; }
        nd.mass = gm; nd.comx = cx; nd.comy = cy; nd.comz = cz;
        t->nodes[idx] = nd;
        return 0;
    }
    /* mid = pos + size * 0.5f; z first, then y in each z half, then x */
    const float half = size * 0.5f;
    const float midx = px + half, midy = py + half, midz = pz + half;
    uint32_t b[9];
    b[0] = s; b[8] = e;
    b[4] = part_point(t->z, b[0], b[8], midz);
    b[2] = part_point(t->y, b[0], b[4], midy);
    b[6] = part_point(t->y, b[4], b[8], midy);
    b[1] = part_point(t->x, b[0], b[2], midx);
    b[3] = part_point(t->x, b[2], b[4], midx);
    b[5] = part_point(t->x, b[4], b[6], midx);
    b[7] = part_point(t->x, b[6], b[8], midx);
    t->nodes[idx] = nd;
    for (int q = 0; q < 8; ++q) {
        if (b[q + 1] > b[q]) {
            const float nx = px + ((q & 1) ? half : 0.0f);
            const float ny = py + ((q & 2) ? half : 0.0f);
            const float nz = pz + ((q & 4) ? half : 0.0f);
            int64_t ci = t3_alloc(t);
            if (ci < 0) return -1;
            if (t3_node(t, ci, nx, ny, nz, half, b[q], b[q + 1]) != 0) return -1;
            nd.child[q] = (uint32_t)ci;
        }
    }
    /* subGrids[x][y][z] loops: q = x + 2y + 4z in the order 0,4,2,6,1,5,3,7 */
    static const int order[8] = {0, 4, 2, 6, 1, 5, 3, 7};
    float gm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    uint32_t next = 0;
    for (int k = 0; k < 8; ++k) {
        const uint32_t c = nd.child[order[k]];
        if (c == U32MAX) continue;
        const geo3_node *ch = &t->nodes[c];
        gm = gm + ch->mass;
        cx = cx + ch->comx * ch->mass;
        cy = cy + ch->comy * ch->mass;
        cz = cz + ch->comz * ch->mass;
        next += 1 + ch->next;
    }
    if (gm > 0) { cx = cx / gm; cy = cy / gm; cz = cz / gm; }
    nd.mass = gm; nd.comx = cx; nd.comy = cy; nd.comz = cz; nd.next = next;
    t->nodes[idx] = nd;
    return 0;
}

static int64_t tree3_internal(int32_t n, const float *x, const float *y, const float *z, const float *m,
                              const float bb[4], uint32_t leaf_cap, geo3_node **out) {
    tree3 t;
    t.leaf_cap = leaf_cap;
    t.cap = 8 * (int64_t)n + 16; /* Octree::root reserves 8N (quadtree.cpp:279) */
    t.count = 0;
    t.nodes = (geo3_node *)malloc(sizeof(geo3_node) * (size_t)t.cap);
    t.x = x; t.y = y; t.z = z; t.m = m;
    if (!t.nodes) return -1;
    int64_t r = t3_alloc(&t);
    if (r < 0 || t3_node(&t, r, bb[0], bb[1], bb[2], bb[3], 0, (uint32_t)n) != 0) {
        free(t.nodes);
        return -1;
    }
    *out = t.nodes;
    return t.count;
}

int32_t geo3_build_tree(int32_t n, const float *x, const float *y, const float *z, const float *m,
                        const float bb[4], uint32_t leaf_cap, geo3_node *nodes, int32_t cap) {
    geo3_node *tmp = NULL;
    int64_t c = tree3_internal(n, x, y, z, m, bb, leaf_cap, &tmp);
    if (c < 0) return -1;
    int64_t k = c < cap ? c : cap;
    if (nodes && k > 0) memcpy(nodes, tmp, sizeof(geo3_node) * (size_t)k);
    free(tmp);
    return (int32_t)c;
}

/* ------------------------------------------------------------ gravity --- */
/* d -= size * ((d > half) - (d < -half)), physics3D.cpp:160-164 */
static inline float wrap(float d, float size, float half) {
    return d - size * (float)((d > half) - (d < -half));
}

void geo3_bh_accel(int32_t n, const float *x, const float *y, const float *z, const float *m,
                   const geo3_node *nodes, int32_t nn, float G, float theta, float eps,
                   const geo3_domain *dom, float *ax, float *ay, float *az, uint64_t *interactions) {
    const int per = dom && dom->w > 0.0f;
    const float W = per ? dom->w : 0.0f, H = per ? dom->h : 0.0f, D = per ? dom->d : 0.0f;
    const float hw = W * 0.5f, hh = H * 0.5f, hd = D * 0.5f;
    const float thetaSq = theta * theta;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int32_t i = 0; i < n; ++i) {
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        const float pmass = m[i];
        const float xi = x[i], yi = y[i], zi = z[i];
        uint32_t idx = 0;
        uint64_t cnt = 0;
        while (idx < (uint32_t)nn) {
            const geo3_node *g = &nodes[idx];
            const float gmass = g->mass;
            if (gmass <= 0.0f) { idx += g->next + 1; continue; }
            float dx = g->comx - xi, dy = g->comy - yi, dz = g->comz - zi;
            if (per) { dx = wrap(dx, W, hw); dy = wrap(dy, H, hh); dz = wrap(dz, D, hd); }
            /* softening added unsquared (physics3D.cpp:166) */
            const float distSq = dx * dx + dy * dy + dz * dz + eps;
            const int leaf = g->next == 0; /* every subGrids entry empty */
            const float sizeSq = g->size * g->size;
            if (sizeSq < thetaSq * distSq) {
                const float inv = 1.0f / sqrtf(distSq);
                const float inv3 = inv * inv * inv;
                const float fm = G * pmass * gmass * inv3;
                fx = fx + dx * fm;
                fy = fy + dy * fm;
                fz = fz + dz * fm;
                cnt++;
                idx += g->next + 1;
            } else if (leaf) {
                for (uint32_t j = g->start; j < g->end; ++j) {
                    if ((uint32_t)i == j) continue;
                    float px = x[j] - xi, py = y[j] - yi, pz = z[j] - zi;
                    if (per) { px = wrap(px, W, hw); py = wrap(py, H, hh); pz = wrap(pz, D, hd); }
                    const float d2 = px * px + py * py + pz * pz + eps;
                    const float inv = 1.0f / sqrtf(d2);
                    const float inv3 = inv * inv * inv;
                    const float fm = G * pmass * m[j] * inv3;
                    fx = fx + px * fm;
                    fy = fy + py * fm;
                    fz = fz + pz * fm;
                    cnt++;
                }
                idx += g->next + 1;
            } else {
                ++idx;
            }
        }
        ax[i] = fx / m[i];
        ay[i] = fy / m[i];
        az[i] = fz / m[i];
        total += cnt;
    }
    if (interactions) *interactions = total;
}

void geo3_direct_accel(int32_t n, const float *x, const float *y, const float *z, const float *m,
                       float G, float eps, const geo3_domain *dom, float *ax, float *ay, float *az) {
    const int per = dom && dom->w > 0.0f;
    const float W = per ? dom->w : 0.0f, H = per ? dom->h : 0.0f, D = per ? dom->d : 0.0f;
    const float hw = W * 0.5f, hh = H * 0.5f, hd = D * 0.5f;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 0; i < n; ++i) {
        const float pix = x[i], piy = y[i], piz = z[i];
        float tx = 0.0f, ty = 0.0f, tz = 0.0f;
        for (int32_t j = 0; j < n; ++j) {
            float dx = x[j] - pix, dy = y[j] - piy, dz = z[j] - piz;
            if (per) { dx = wrap(dx, W, hw); dy = wrap(dy, H, hh); dz = wrap(dz, D, hd); }
            const float distSq = dx * dx + dy * dy + dz * dz + eps;
            const float inv = 1.0f / sqrtf(distSq);
            const float inv3 = inv * inv * inv;
            const float factor = G * m[j] * inv3;
            tx = tx + dx * factor;
            ty = ty + dy * factor;
            tz = tz + dz * factor;
        }
        ax[i] = tx;
        ay[i] = ty;
        az[i] = tz;
    }
}

void geo3_direct_fp64(int32_t n, const float *x, const float *y, const float *z, const float *m,
                      double G, double eps, int32_t nt, const uint32_t *targets,
                      double *ax, double *ay, double *az) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t t = 0; t < nt; ++t) {
        const uint32_t i = targets ? targets[t] : (uint32_t)t;
        const double xi = x[i], yi = y[i], zi = z[i];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int32_t j = 0; j < n; ++j) {
            if ((uint32_t)j == i) continue;
            const double dx = (double)x[j] - xi, dy = (double)y[j] - yi, dz = (double)z[j] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz + eps;
            const double inv = 1.0 / sqrt(r2);
            const double f = (double)m[j] * inv * inv * inv;
            sx += dx * f;
            sy += dy * f;
            sz += dz * f;
        }
        ax[t] = G * sx;
        ay[t] = G * sy;
        az[t] = G * sz;
    }
}

/* ---------------------------------------------------------- integrator --- */
/* Physics3D::integrateStart3D, physics3D.cpp:740-804 */
void geo3_integrate_start(geo3_state *s, float dt, int32_t sph_clamp, float sph_max_vel,
                          const geo3_domain *dom) {
    const float halfDt = dt * 0.5f;
    const float maxVSq = sph_max_vel * sph_max_vel;
    const int per = dom && dom->w > 0.0f;
    const float hw = per ? dom->w * 0.5f : 0.0f, hh = per ? dom->h * 0.5f : 0.0f;
    const float hd = per ? dom->d * 0.5f : 0.0f;
    for (int32_t i = 0; i < s->n; ++i) {
        if (s->flags[i] & GEO_PINNED) continue;
        float pvx = s->vx[i], pvy = s->vy[i], pvz = s->vz[i];
        float vx = s->vx[i] + s->ax[i] * halfDt;
        float vy = s->vy[i] + s->ay[i] * halfDt;
        float vz = s->vz[i] + s->az[i] * halfDt;
        if (sph_clamp) {
            const float vSq = vx * vx + vy * vy + vz * vz;
            if (vSq > maxVSq) {
                const float pSq = pvx * pvx + pvy * pvy + pvz * pvz;
                if (pSq > 0.00001f) {
                    const float inv = sph_max_vel / sqrtf(pSq);
                    pvx = pvx * inv; pvy = pvy * inv; pvz = pvz * inv;
                }
                const float inv = sph_max_vel / sqrtf(vSq);
                vx = vx * inv; vy = vy * inv; vz = vz * inv;
            }
        }
        s->pvx[i] = pvx; s->pvy[i] = pvy; s->pvz[i] = pvz;
        s->vx[i] = vx; s->vy[i] = vy; s->vz[i] = vz;
        float px = s->x[i] + vx * dt, py = s->y[i] + vy * dt, pz = s->z[i] + vz * dt;
        if (per) { /* physics3D.cpp:782-800: the cube [-half, half) */
            if (px < -hw) px += dom->w;
            else if (px >= hw) px -= dom->w;
            if (py < -hh) py += dom->h;
            else if (py >= hh) py -= dom->h;
            if (pz < -hd) pz += dom->d;
            else if (pz >= hd) pz -= dom->d;
        }
        s->x[i] = px; s->y[i] = py; s->z[i] = pz;
    }
}

/* Physics3D::integrateEnd3D, physics3D.cpp:1158-1168 */
void geo3_integrate_end(geo3_state *s, float dt) {
    const float h = dt * 0.5f;
    for (int32_t i = 0; i < s->n; ++i) {
        if (s->flags[i] & GEO_PINNED) continue;
        s->vx[i] = s->vx[i] + s->ax[i] * h;
        s->vy[i] = s->vy[i] + s->ay[i] * h;
        s->vz[i] = s->vz[i] + s->az[i] * h;
    }
}

static void permute3(geo3_state *s, const uint32_t *perm) {
    const int32_t n = s->n;
    uint32_t *t = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    void *fs[] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az,
                  s->pvx, s->pvy, s->pvz, s->m, s->flags, s->id,
                  s->sph_mass, s->rest_dens, s->stiff, s->visc, s->cohesion, s->dens,
                  s->press, s->press_tmp, s->pred_dens, s->ppx, s->ppy, s->ppz, s->pfx, s->pfy, s->pfz};
    for (size_t k = 0; k < sizeof(fs) / sizeof(fs[0]); ++k) {
        uint32_t *a = (uint32_t *)fs[k];
        if (!a) continue;
        for (int32_t i = 0; i < n; ++i) t[i] = a[perm[i]];
        memcpy(a, t, sizeof(uint32_t) * (size_t)n);
    }
    free(t);
}

/* -------------------------------------------------------------- SPH3D --- */
#define PI_F3 3.14159265358979323846f /* raylib PI (a float) */

/* SPH3D.h:27-53.  pow with float arguments resolves to std::pow(float, float)
 * once <math.h> is included (raymath.h does), so the 3D kernels compute in
 * float: powf. */
void geo3_sph_constants(float h, float out[3]) {
    out[0] = (2.0f * PI_F3 * powf(h, 5.0f)) / 15.0f; /* smoothingKernel volume */
    out[1] = -45.0f / (PI_F3 * powf(h, 6.0f));        /* spikyKernelDerivative scale */
    out[2] = 45.0f / (PI_F3 * powf(h, 6.0f));         /* smoothingKernelLaplacian scale */
}

/* SPH3D::smoothingKernelCohesion, SPH3D.h:48-53 */
static inline float coh3(float r, float h) {
    if (r >= h) return 0.0f;
    float q = r / h;
    return (1.0f - q) * (0.5f - q) * (0.5f - q) * 30.0f / (PI_F3 * h * h * h);
}

typedef struct { int32_t cx, cy, cz; uint32_t i; } cell3_entry;

static int cell3_cmp(const void *a, const void *b) {
    const cell3_entry *p = (const cell3_entry *)a, *q = (const cell3_entry *)b;
    if (p->cz != q->cz) return p->cz < q->cz ? -1 : 1;
    if (p->cy != q->cy) return p->cy < q->cy ? -1 : 1;
    if (p->cx != q->cx) return p->cx < q->cx ? -1 : 1;
    return p->i < q->i ? -1 : (p->i > q->i);
}

/* NeighborSearch3DV2AVX2::posToCellCoord, neighborSearch.h:1005-1010 */
static inline int32_t cell3(float p, float cell) { return (int32_t)floorf(p / cell); }

typedef struct { cell3_entry *e; int32_t n; float cell; } cell3_grid;

/* newGridAVX2 (neighborSearch.h:1017-1107) restated as a (cz, cy, cx, index)
 * ordered table; its candidate SET equals the hashed table's, whose queries
 * compare exact cell coordinates. */
static cell3_grid grid3_build(int32_t n, const float *x, const float *y, const float *z, float cell) {
    cell3_grid g;
    g.n = n;
    g.cell = cell;
    g.e = (cell3_entry *)malloc(sizeof(cell3_entry) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; ++i) {
        g.e[i].cx = cell3(x[i], cell);
        g.e[i].cy = cell3(y[i], cell);
        g.e[i].cz = cell3(z[i], cell);
        g.e[i].i = (uint32_t)i;
    }
    qsort(g.e, (size_t)n, sizeof(cell3_entry), cell3_cmp);
    return g;
}

static int32_t grid3_lower(const cell3_grid *g, int32_t cx, int32_t cy, int32_t cz) {
    int32_t lo = 0, hi = g->n;
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        const cell3_entry *e = &g->e[mid];
        int less = e->cz < cz || (e->cz == cz && (e->cy < cy || (e->cy == cy && e->cx < cx)));
        if (less) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* queryNeighborsAVX2 (neighborSearch.h:1109-1173): dz, dy, dx from -1 to 1,
 * every particle whose cell matches; no distance filter, self included */
static int32_t grid3_query(const cell3_grid *g, float qx, float qy, float qz, uint32_t *buf, int32_t cap) {
    const int32_t cx = cell3(qx, g->cell), cy = cell3(qy, g->cell), cz = cell3(qz, g->cell);
    int32_t k = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int32_t nx = cx + dx, ny = cy + dy, nz = cz + dz;
                int32_t p = grid3_lower(g, nx, ny, nz);
                while (p < g->n && g->e[p].cx == nx && g->e[p].cy == ny && g->e[p].cz == nz) {
                    if (k < cap) buf[k] = g->e[p].i;
                    ++k;
                    ++p;
                }
            }
    return k;
}

static int u32_cmp3(const void *a, const void *b) {
    uint32_t p = *(const uint32_t *)a, q = *(const uint32_t *)b;
    return p < q ? -1 : (p > q);
}

int64_t geo3_sph_candidates(int32_t n, const float *x, const float *y, const float *z, int32_t nq,
                            const float *qx, const float *qy, const float *qz, float cell,
                            int64_t *off, uint32_t *idx, int64_t cap) {
    cell3_grid g = grid3_build(n, x, y, z, cell);
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    int64_t total = 0;
    for (int32_t q = 0; q < nq; ++q) {
        off[q] = total;
        int32_t k = grid3_query(&g, qx[q], qy[q], qz[q], buf, n);
        qsort(buf, (size_t)k, sizeof(uint32_t), u32_cmp3);
        if (idx && total + k <= cap) memcpy(idx + total, buf, sizeof(uint32_t) * (size_t)k);
        total += k;
    }
    off[nq] = total;
    free(buf);
    free(g.e);
    return total <= cap ? total : -1;
}

/* SPH3D::computeViscCohesionForces (SPH3D.cpp:98-170) and SPH3D::PCISPH
 * (SPH3D.cpp:172-308) with maxIter = 1.  The pair sums are applied in a fixed
 * order (query order, particle by particle) where the reference uses OpenMP
 * atomics. */
int32_t geo3_sph_pcisph(geo3_state *s, const geo3_params *p) {
    const int32_t n = s->n;
    const float h = p->sph_h, h2 = h * h;
    float kc[3];
    geo3_sph_constants(h, kc);
    const float volume = kc[0], spiky = kc[1], lapS = kc[2];
    const float coeff = 0.5f * p->dt * p->dt; /* predictionCoeff, SPH3D.cpp:176 */
    const size_t nn = (size_t)(n > 0 ? n : 1);
    float *fx = (float *)calloc(nn, sizeof(float));
    float *fy = (float *)calloc(nn, sizeof(float));
    float *fz = (float *)calloc(nn, sizeof(float));
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * nn);
    cell3_grid g = grid3_build(n, s->x, s->y, s->z, p->sph_cell);
    const uint32_t *fl = s->flags;
#define IS_SPH(k) ((fl[k] & GEO_SPH) != 0)
#define DRAWN(k) ((fl[k] & GEO_DRAWN) != 0)
#define PINNED(k) ((fl[k] & GEO_PINNED) != 0)
    /* viscosity + cohesion, SPH3D.cpp:98-170 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || PINNED(i) || DRAWN(i)) continue;
        int32_t k = grid3_query(&g, s->x[i], s->y[i], s->z[i], buf, n);
        for (int32_t t = 0; t < k; ++t) {
            const uint32_t j = buf[t];
            if (!IS_SPH(j) || DRAWN(j)) continue;
            if ((int32_t)j == i) continue;
            const float dx = s->x[j] - s->x[i], dy = s->y[j] - s->y[i], dz = s->z[j] - s->z[i];
            const float rSq = dx * dx + dy * dy + dz * dz;
            if (rSq >= h2) continue;
            const float r = sqrtf(rSq < 1e-6f ? 1e-6f : rSq); /* std::max(rSq, 1e-6f) */
            const float nrx = dx / r, nry = dy / r, nrz = dz / r;
            const float mJ = s->sph_mass[j] * p->sph_unit_mass;
            const float lapW = r >= h ? 0.0f : lapS * (h - r); /* smoothingKernelLaplacian */
            const float dj = s->dens[j] < 0.001f ? 0.001f : s->dens[j];
            const float vfx = p->viscosity * s->visc[j] * mJ / dj * lapW * (s->vx[j] - s->vx[i]);
            const float vfy = p->viscosity * s->visc[j] * mJ / dj * lapW * (s->vy[j] - s->vy[i]);
            const float vfz = p->viscosity * s->visc[j] * mJ / dj * lapW * (s->vz[j] - s->vz[i]);
            const float cohCoef = p->cohesion_coef * s->cohesion[i];
            const float cf = coh3(r, h);
            const float cfx = cohCoef * mJ * cf * nrx, cfy = cohCoef * mJ * cf * nry, cfz = cohCoef * mJ * cf * nrz;
            fx[i] = fx[i] + (vfx + cfx);
            fy[i] = fy[i] + (vfy + cfy);
            fz[i] = fz[i] + (vfz + cfz);
            fx[j] = fx[j] - (vfx + cfx);
            fy[j] = fy[j] - (vfy + cfy);
            fz[j] = fz[j] - (vfz + cfz);
        }
    }
    for (int32_t i = 0; i < n; ++i) { /* SPH3D.cpp:182-185 */
        s->press[i] = 0.0f;
        s->pfx[i] = 0.0f; s->pfy[i] = 0.0f; s->pfz[i] = 0.0f;
    }
    /* predicted positions, SPH3D.cpp:190-198 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || DRAWN(i)) continue;
        s->ppx[i] = s->x[i] + (fx[i] / s->sph_mass[i]) * coeff;
        s->ppy[i] = s->y[i] + (fy[i] / s->sph_mass[i]) * coeff;
        s->ppz[i] = s->z[i] + (fz[i] / s->sph_mass[i]) * coeff;
    }
    const float m2 = p->sph_unit_mass * 2.0f; /* myVar.mass * 2.0f, SPH3D.cpp:225, 262 */
    /* predicted density and pressure, SPH3D.cpp:200-236 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || DRAWN(i)) continue;
        float pd = 0.0f;
        int32_t k = grid3_query(&g, s->ppx[i], s->ppy[i], s->ppz[i], buf, n);
        for (int32_t t = 0; t < k; ++t) {
            const uint32_t j = buf[t];
            if (!IS_SPH(j) || DRAWN(j)) continue;
            const float dx = s->ppx[i] - s->ppx[j], dy = s->ppy[i] - s->ppy[j], dz = s->ppz[i] - s->ppz[j];
            const float rrSq = dx * dx + dy * dy + dz * dz;
            if (rrSq >= h * h) continue;
            const float rr = sqrtf(rrSq);
            const float mJ = s->sph_mass[j] * m2;
            const float rho0 = 0.5f * (s->rest_dens[i] + s->rest_dens[j]);
            const float w = rr >= h ? 0.0f : powf(h - rr, 2.0f) / volume; /* smoothingKernel */
            pd = pd + mJ * w / rho0;
        }
        s->pred_dens[i] = pd;
        const float err = pd - s->rest_dens[i];
        float pt = p->delta * err;
        if (pt < 0.0f) pt = 0.0f;
        s->press_tmp[i] = pt;
        s->press[i] = s->press[i] + pt * s->stiff[i] * p->stiff_mult;
    }
    /* pressure force, SPH3D.cpp:238-291 */
    for (int32_t i = 0; i < n; ++i) {
        if (!IS_SPH(i) || DRAWN(i)) continue;
        int32_t k = grid3_query(&g, s->ppx[i], s->ppy[i], s->ppz[i], buf, n);
        for (int32_t t = 0; t < k; ++t) {
            const uint32_t j = buf[t];
            if ((int32_t)j == i) continue;
            if (!IS_SPH(j) || DRAWN(j)) continue;
            const float dx = s->ppx[i] - s->ppx[j], dy = s->ppy[i] - s->ppy[j], dz = s->ppz[i] - s->ppz[j];
            const float rr = sqrtf(dx * dx + dy * dy + dz * dz);
            if (rr < 1e-5f || rr >= h) continue;
            const float gradW = spiky * powf(h - rr, 2.0f); /* spikyKernelDerivative */
            const float nx = dx / rr, ny = dy / rr, nz = dz / rr;
            const float avgP = 0.5f * (s->press[i] + s->press[j]);
            const float avgD = 0.5f * (s->pred_dens[i] + s->pred_dens[j]);
            float mag = -(s->sph_mass[i] * m2 + s->sph_mass[j] * m2) * avgP / (avgD < 0.01f ? 0.01f : avgD);
            const float smax = s->sph_mass[i] < s->sph_mass[j] ? s->sph_mass[j] : s->sph_mass[i];
            const float smin = s->sph_mass[j] < s->sph_mass[i] ? s->sph_mass[j] : s->sph_mass[i];
            const float ratio = smax / smin;
            float sc = 8.0f / ratio;
            if (1.0f < sc) sc = 1.0f;
            mag = mag * sc;
            const float pfx = mag * gradW * nx, pfy = mag * gradW * ny, pfz = mag * gradW * nz;
            fx[i] = fx[i] + pfx; fy[i] = fy[i] + pfy; fz[i] = fz[i] + pfz;
            fx[j] = fx[j] - pfx; fy[j] = fy[j] - pfy; fz[j] = fz[j] - pfz;
        }
    }
    /* pressF and acceleration, SPH3D.cpp:296-307 */
    for (int32_t i = 0; i < n; ++i) {
        s->pfx[i] = fx[i] / s->sph_mass[i];
        s->pfy[i] = fy[i] / s->sph_mass[i];
        s->pfz[i] = fz[i] / s->sph_mass[i];
        if (!PINNED(i)) {
            s->ax[i] = s->ax[i] + s->pfx[i];
            s->ay[i] = s->ay[i] + s->pfy[i];
            s->az[i] = s->az[i] + s->pfz[i];
        }
    }
#undef IS_SPH
#undef DRAWN
#undef PINNED
    free(g.e);
    free(buf);
    free(fx);
    free(fy);
    free(fz);
    return 0;
}

/* SPH3D::groundModeBoundary, SPH3D.cpp:310-366: walls at +-domain/2 */
void geo3_sph_ground(geo3_state *s, float radius, float w, float h, float d, float vertical_gravity,
                     float friction) {
    const float damp = -0.1f; /* SPH3D::boundDamping (SPH3D.h:14) */
    for (int32_t i = 0; i < s->n; ++i) {
        if (s->flags[i] & GEO_PINNED) continue;
        float x = s->x[i], y = s->y[i], z = s->z[i], vx = s->vx[i], vy = s->vy[i], vz = s->vz[i];
        s->ay[i] = s->ay[i] - vertical_gravity;
        int hx = 0, hy = 0, hz = 0;
        if (x - radius < -w / 2.0f) { vx *= damp; x = -w / 2.0f + radius; hx = 1; }
        if (x + radius > w / 2.0f) { vx *= damp; x = w / 2.0f - radius; hx = 1; }
        if (y - radius < -h / 2.0f) { vy *= damp; y = -h / 2.0f + radius; hy = 1; }
        if (y + radius > h / 2.0f) { vy *= damp; y = h / 2.0f - radius; hy = 1; }
        if (z - radius < -d / 2.0f) { vz *= damp; z = -d / 2.0f + radius; hz = 1; }
        if (z + radius > d / 2.0f) { vz *= damp; z = d / 2.0f - radius; hz = 1; }
        if (hx) { vy *= 1.0f - friction; vz *= 1.0f - friction; }
        if (hy) { vx *= 1.0f - friction; vz *= 1.0f - friction; }
        if (hz) { vx *= 1.0f - friction; vy *= 1.0f - friction; }
        s->x[i] = x; s->y[i] = y; s->z[i] = z;
        s->vx[i] = vx; s->vy[i] = vy; s->vz[i] = vz;
    }
}

/* mode3D's physics frame, GE/src/globalLogic.cpp:1580-1719 */
int32_t geo3_step(geo3_state *s, const geo3_params *p, int32_t nsteps) {
    const int32_t n = s->n;
    if (n <= 0) return 0;
    geo3_domain dom = {0.0f, 0.0f, 0.0f};
    if (p->periodic) { dom.w = p->domain_w; dom.h = p->domain_h; dom.d = p->domain_d; }
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint32_t *perm = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int32_t step = 0; step < nsteps; ++step) {
        const float dt0 = step == 0 ? p->dt_start : p->dt;
        geo3_node *nodes = NULL;
        int64_t nn = 0;
        if (dt0 != 0.0f) {
            geo3_integrate_start(s, dt0, p->sph_clamp || p->sph, p->sph_max_vel, &dom);
            float bb[4];
            geo3_bbox(n, s->x, s->y, s->z, bb);
            geo3_morton_keys(n, s->x, s->y, s->z, bb, keys);
            geo_sort_perm(n, keys, perm);
            permute3(s, perm);
            if (!p->naive) {
                nn = tree3_internal(n, s->x, s->y, s->z, s->m, bb, p->leaf_cap, &nodes);
                if (nn < 0) { free(keys); free(perm); return -2; }
            }
        }
        if (p->dt != 0.0f) {
            for (int32_t i = 0; i < n; ++i) { s->ax[i] = 0.0f; s->ay[i] = 0.0f; s->az[i] = 0.0f; }
            if (p->G != 0.0f) {
                if (p->naive || n < p->direct_below || !nodes)
                    geo3_direct_accel(n, s->x, s->y, s->z, s->m, p->G, p->softening, &dom, s->ax, s->ay, s->az);
                else
                    geo3_bh_accel(n, s->x, s->y, s->z, s->m, nodes, (int32_t)nn, p->G, p->theta, p->softening,
                                  &dom, s->ax, s->ay, s->az, NULL);
            }
            if (p->sph) { /* sph3D.pcisphSolver, globalLogic.cpp:1707-1709 */
                geo3_sph_pcisph(s, p);
                if (p->sph_ground)
                    geo3_sph_ground(s, p->sph_h, p->domain_w, p->domain_h, p->domain_d, p->vertical_gravity,
                                    p->boundary_friction);
            }
            geo3_integrate_end(s, p->dt);
        }
        free(nodes);
    }
    free(keys);
    free(perm);
    return 0;
}
