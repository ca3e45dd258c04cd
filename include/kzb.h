/*
 * kzb.h - batched worlds: thousands of independent PhysicsWorlds of uneven
 * size stepped together on one GPU (libkizuri_amd.so).
 *
 * Every world is a reference PhysicsWorld of its own.  Each has its own
 * bodies, shapes, ground plane, broadphase and warm-start cache, and StepFixed
 * (src/World/PhysicsWorld.cpp:245-329) semantics: a world's trajectory is
 * bit-identical to stepping it alone through kzw/kzg.  All worlds share one
 * kz_world_desc (gravity, iterations, solver switches).
 *
 * Per step:
 *   - the host updates each world's dynamic AABB tree (host threads) and
 *     uploads the concatenated pair list;
 *   - one captured HIP graph runs the whole contact path of every world:
 *     narrowphase, constraint compaction, constraint build, islands, the
 *     reference's first-fit batches, the level schedule, prepare, the
 *     persistent solve kernel (one workgroup per world: warm start,
 *     velocity and bias sweeps), write-back, cache store and integration;
 *   - the body states are copied back to the host after the graph.
 *
 * The launch configuration (waves for the iterative narrowphase, threads
 * per world in the solve) never changes results; kzb_autotune times
 * candidates on live steps and keeps the fastest.
 *
 * Limits of the batched path (a world beyond them makes kzb_step fail with
 * KZG_ERR_UNSUPPORTED): 1024 bodies, 2048 first-fit batches and 4096
 * levels per world.
 */
#ifndef KZB_H
#define KZB_H

#include <stdint.h>
#include "kz_scene.h"
#include "kzg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kzb_batch kzb_batch;

typedef struct kzb_config {
    uint32_t narrowphase_waves;  /* workgroups (one wave each) of the iterative narrowphase; 0 = 16 per CU */
    uint32_t solve_threads;      /* threads per world in the solve kernel: 64..1024, power of two */
    uint32_t broadphase_threads; /* host threads for the per-world broadphase; 0 = hardware */
    uint32_t flags;              /* KZB_DIRECT_LAUNCH */
} kzb_config;

/* kzb_config.flags: launch the step's kernels one by one instead of replaying
 * the captured graph, with events between the phases (kzb_stats.phase_ms). */
#define KZB_DIRECT_LAUNCH 1u

typedef struct kzb_stats {
    uint32_t num_worlds, num_bodies, num_pairs, num_constraints, num_points;
    uint32_t max_velocity_levels, max_bias_levels;
    uint32_t reserved;
    double broadphase_ms;   /* host: every world's broadphase */
    double device_ms;       /* HIP events around the graph: the device step */
    double step_ms;         /* host clock around kzb_step */
    double device_with_copies_ms;  /* the graph and the copies of the body states to the host */
    double launch_ms;       /* host clock around the graph launch (or the direct launches) */
    /* KZB_DIRECT_LAUNCH only: narrowphase, compaction + constraint build,
     * schedule, integration + prepare, sweeps, write-back to the end */
    float phase_ms[6];
} kzb_stats;

int kzb_create(kzb_batch** out, const kz_world_desc* settings, int device);
void kzb_destroy(kzb_batch* b);
const char* kzb_last_error(const kzb_batch* b);

/* Append a world (shapes with hull points as in kzw_add_shapes, bodies as
 * in kzw_add_bodies, body shape indices local to the world).  Returns the
 * world index (>= 0) or a negative status. */
int kzb_add_world(kzb_batch* b, const kz_shape_desc* shapes, uint32_t num_shapes, const float* hull_points,
                  uint32_t num_hull_points, const kz_body_desc* bodies, uint32_t num_bodies);

/* StepFixed(dt) of every world. */
int kzb_step(kzb_batch* b, float dt);

int kzb_num_worlds(kzb_batch* b);
int kzb_get_states(kzb_batch* b, uint32_t world, kz_body_state* out, uint32_t n);
int kzb_get_stats(kzb_batch* b, kzb_stats* out);
int kzb_set_config(kzb_batch* b, const kzb_config* cfg);
int kzb_get_config(kzb_batch* b, kzb_config* cfg);

/* Step the batch (dt each step) while timing candidate configurations:
 * every candidate runs `repeats` steps, interleaved; the one with the lowest
 * median device time is kept.  The trajectory is the same whatever the
 * configuration.  times[] (optional, one per candidate in the order
 * reported by kzb_autotune_candidates) receives the medians in ms. */
int kzb_autotune(kzb_batch* b, float dt, uint32_t repeats, kzb_config* best, float* times, uint32_t cap);
int kzb_autotune_candidates(kzb_batch* b, kzb_config* out, uint32_t cap);

/* Schedule of the last step for one world (tests): per world-local contact
 * point its velocity level and warm-start/bias level (levels from 1;
 * velocity level 0 = not swept).  Returns the world's point count. */
int kzb_world_schedule(kzb_batch* b, uint32_t world, uint32_t* vel_level, uint32_t* ws_level, uint32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* KZB_H */
