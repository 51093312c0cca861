/*
 * ge_serve.h -- host layer above the C ABI of ge_physics.h: the autotuner of
 * the batched form and the continuous-batching request server.
 *
 * Neither has a counterpart in the reference, which steps one interactive
 * scene per process (GE/src/globalLogic.cpp:1155-1372, driven once per frame
 * from GE/src/main.cpp).  Both are plain host C++ that call only the entry
 * points of ge_physics.h; all device work goes through those.
 */
#ifndef GE_SERVE_H
#define GE_SERVE_H

#include "ge_physics.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- autotuner ------------------------------------------------------------
 * Times ge_step over `frames` frames (after one untimed frame) for every leaf
 * capacity in {8, 16, 32} and walk block size in {64, 128, 256}, plus the
 * reference leaf rule at block 256, on the particles the context holds now.
 * The context's state is saved to host memory first and restored after every
 * candidate and at the end, so tuning leaves the simulation unchanged.
 * Leaf capacities change the tree (not the opening test), so every candidate
 * opens at least the nodes the reference tree opens at the same node; the
 * choice is by time alone. */
typedef struct {
    uint32_t leaf_capacity;   /* chosen leaf capacity */
    uint32_t walk_block;      /* chosen walk block size */
    float ms_per_frame;       /* time of the chosen pair */
    float reference_ms;       /* reference leaf rule, block 256 */
    float table_ms[3][3];     /* [leaf 8/16/32][block 64/128/256] */
} ge_tune_result;

int ge_autotune(ge_ctx *ctx, const ge_step_params *base, int32_t frames, ge_tune_result *out);

/* ---- continuous-batching server ------------------------------------------
 * Gravity-only requests share one batched context (ge_upload_systems /
 * ge_systems_update): every round steps all resident systems together for
 * min(round_frames, fewest remaining steps) frames, then finished systems are
 * read back and dropped and waiting requests join, while the others stay
 * resident in HBM.  Gravity + SPH requests each take a single-system SPH
 * context of their own (one stream each) and run concurrently with the
 * batch, round_frames frames per chunk.  Batch rounds and SPH chunks run
 * asynchronously: the loop wakes on whichever finishes first, so an SPH
 * context gets its next chunk as soon as its last one ends (the batch
 * changes membership only between its rounds).  Admission is first come,
 * first served, within an HBM budget: the
 * contexts are sized from the budget at creation, and a request is admitted
 * when its particles fit the free capacity (and, for SPH, a free context).
 * With batching = 0 the server is the plain form: one request on the device
 * at a time, in arrival order.
 *
 * Every request runs with the server's parameter set for its kind, with
 * dt_start = dt, so its result is the same as stepping it alone:
 * ge_upload + ge_step(steps) on a fresh context. */
typedef struct ge_server ge_server;

typedef struct {
    int32_t device;
    uint64_t hbm_budget_bytes;   /* device memory all contexts may hold */
    uint32_t batch_capacity;     /* upper bound on particles of the batch */
    uint32_t max_systems;        /* resident systems in the batch (<= 2048) */
    uint32_t sph_slots;          /* SPH contexts (0: SPH requests rejected) */
    uint32_t sph_capacity;       /* particles per SPH context */
    uint32_t round_frames;       /* frames per round between admission points */
    int32_t batching;            /* 1: continuous batching; 0: plain form */
    int32_t use_graph;           /* replay frames through hipGraphs */
    uint32_t reserved[4];
} ge_server_config;

typedef struct {
    uint32_t n;                  /* particles */
    int32_t steps;               /* frames to run */
    int32_t sph;                 /* 0: gravity only; 1: gravity + SPH */
    double arrival_s;            /* arrival, seconds after ge_server_run starts */
    ge_particles in;             /* host arrays read at admission */
    ge_particles out;            /* host arrays written at completion (NULL fields skipped),
                                    particles in device (Morton) order with their ids */
} ge_request;

typedef struct {
    double arrival_s, admit_s, done_s;  /* seconds after the run started */
    int32_t status;                     /* GE_OK or the failing call's status */
} ge_request_result;

typedef struct {
    double wall_s;               /* first arrival to last completion */
    uint64_t particle_updates;   /* sum over requests of n * steps */
    uint32_t completed;
    double p50_s, p99_s, mean_s; /* time to completion (done - arrival) */
    double solves_per_s;         /* completed / wall_s */
    double updates_per_s;        /* particle_updates / wall_s */
    uint32_t max_resident;       /* most requests on the device at once */
    uint64_t device_bytes;       /* memory of all contexts (<= budget) */
    uint64_t rounds;             /* batch rounds launched */
} ge_server_stats;

void ge_server_default_config(ge_server_config *cfg);
/* gravity: parameters of gravity-only requests; sph: of SPH requests (NULL
 * when sph_slots = 0).  Fails with GE_E_NOMEM when the contexts do not fit
 * the budget. */
int ge_server_create(ge_server **out, const ge_server_config *cfg, const ge_step_params *gravity,
                     const ge_step_params *sph);
int ge_server_destroy(ge_server *srv);
const char *ge_server_last_error(const ge_server *srv);
/* runs nreq requests (any arrival order) to completion in real time.  Every
 * run starts from idle contexts and an empty batch, whatever the previous run
 * left: a run that returned an error (its unfinished requests keep done_s =
 * -1) does not affect the next one. */
int ge_server_run(ge_server *srv, uint32_t nreq, const ge_request *reqs, ge_request_result *res,
                  ge_server_stats *stats);
/* A bounded run: ge_server_run stops with GE_E_CAPACITY ("round limit
 * reached") once it has launched max_rounds batch rounds, after waiting for
 * the work still queued; 0 (the default) removes the limit.  Lets a caller
 * drain a server on a deadline; requests not finished by then keep
 * done_s = -1. */
int ge_server_set_round_limit(ge_server *srv, uint64_t max_rounds);

#ifdef __cplusplus
}
#endif
#endif /* GE_SERVE_H */
