/*
 * kzg.h - C-ABI of the MI355X contact path (libkizuri_amd.so).
 *
 * The entry points replace, on the device, steps 2, 3 and 6 of the
 * reference's PhysicsWorld::StepFixed (src/World/PhysicsWorld.cpp:245-329)
 * together with the velocity and position integration around them:
 *
 *   kzg_narrowphase          PhysicsWorld::Collide            (PhysicsWorld.cpp:350-400)
 *                            -> NarrowPhase::Collide          (src/Collision/CollideShape.cpp:480-592)
 *   kzg_build_constraints    PhysicsWorld::BuildContactConstraints (PhysicsWorld.cpp:454-535)
 *   kzg_integrate_velocities PhysicsWorld::IntegrateVelocities     (PhysicsWorld.cpp:757-783)
 *   kzg_solve                PhysicsWorld::SolveContacts           (PhysicsWorld.cpp:540-582):
 *                            Prepare, BuildIslands, SolveIsland (warm start, SIMD
 *                            velocity batches, split-impulse bias), cache store
 *   kzg_integrate_positions  IntegratePositions + pseudo-velocity integration +
 *                            UpdateBodyInertias (PhysicsWorld.cpp:785-802, 275-289, 854-861)
 *   kzg_collide_batch        NarrowPhase::Collide over a batch of (shape, transform) pairs
 *
 * Conventions: plain pointers and sizes, host arrays are borrowed for the
 * duration of a call, outputs go to caller-owned buffers sized with
 * kzg_query_sizes.  Every function returns KZG_OK (0) or a negative status;
 * kzg_last_error gives the text.  There is no CPU fallback: without a usable
 * GPU kzg_create fails with KZG_ERR_DEVICE.  A context is used from one thread.
 */
#ifndef KZG_H
#define KZG_H

#include <stdint.h>
#include "kz_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

enum kzg_status {
    KZG_OK = 0,
    KZG_ERR_ARGUMENT = -1,    /* bad pointer, size or index */
    KZG_ERR_DEVICE = -2,      /* HIP runtime error, no device */
    KZG_ERR_UNSUPPORTED = -3, /* a shape or setting outside the device path */
    KZG_ERR_STATE = -4        /* call out of order */
};

typedef struct kzg_ctx kzg_ctx;

/* One body as uploaded to the device: BodySettings + the mass data that
 * BodyManager::RecomputeMassProperties derives (src/Body/BodyManager.cpp:63-99). */
typedef struct kzg_body {
    uint32_t packed_id;          /* BodyID::Packed() */
    uint32_t shape;              /* index into the uploaded shape table */
    uint32_t motion_type;        /* kz_motion_type */
    uint32_t is_sensor;
    float position[3];
    float rotation[4];
    float linear_velocity[3];
    float angular_velocity[3];
    float inverse_mass;
    float inverse_inertia_local[9]; /* column-major */
    float friction, restitution, linear_damping, angular_damping;
    float gravity_factor, max_linear_velocity, max_angular_velocity;
} kzg_body;

/* Built convex hulls: per shape the range of its vertices and triangles in
 * the pools (ConvexHullShape::Create output, src/Geometry/Shapes.cpp:331-471). */
typedef struct kzg_hull_range {
    uint32_t vert_offset, vert_count, tri_offset, tri_count;
} kzg_hull_range;

/* Per-step phase timings measured with HIP events on the context's stream
 * (milliseconds): narrowphase, build (incl. the host schedule), integrate
 * velocities, solve, integrate positions. */
typedef struct kzg_timings {
    float narrowphase_ms, build_ms, integrate_velocities_ms, solve_ms, integrate_positions_ms;
    float solve_kernel_ms;   /* the persistent sweep kernel alone */
    uint32_t velocity_levels, bias_levels, groups, islands;
} kzg_timings;

/* Create a context on HIP device `device` for a world with settings `world`. */
int kzg_create(kzg_ctx** out, const kz_world_desc* world, int device);
void kzg_destroy(kzg_ctx* ctx);
const char* kzg_last_error(const kzg_ctx* ctx);

/* Shapes: types/params from kz_shape_desc (hull_offset/hull_count ignored),
 * built hull geometry in the pools.  Replaces the shape table. */
int kzg_upload_shapes(kzg_ctx* ctx, const kz_shape_desc* shapes, uint32_t num_shapes,
                      const kzg_hull_range* hulls, const float* hull_verts, uint32_t num_verts,
                      const uint32_t* hull_tris, uint32_t num_tris);

/* Bodies, slot order.  Replaces the body table (the warm-start cache is kept). */
int kzg_upload_bodies(kzg_ctx* ctx, const kzg_body* bodies, uint32_t num_bodies);

/* Overwrite position, rotation and velocities of every body. */
int kzg_upload_states(kzg_ctx* ctx, const kz_body_state* states, uint32_t num_bodies);

/* Step 2: pairs = (slot A, slot B) per broadphase pair, in broadphase order. */
int kzg_narrowphase(kzg_ctx* ctx, const uint32_t* pairs, uint32_t num_pairs);

/* Step 3.  warm_start_mode: 0 = manifold cache (reference default),
 * 1 = cold start (the reference's KZ_NOWS switch). */
int kzg_build_constraints(kzg_ctx* ctx, int warm_start_mode);
int kzg_integrate_velocities(kzg_ctx* ctx, float dt);
/* Step 6. */
int kzg_solve(kzg_ctx* ctx, float dt);
int kzg_integrate_positions(kzg_ctx* ctx, float dt);

/* All of the above in StepFixed order. */
int kzg_step(kzg_ctx* ctx, const uint32_t* pairs, uint32_t num_pairs, float dt);

int kzg_query_sizes(kzg_ctx* ctx, uint32_t* num_pairs, uint32_t* num_constraints, uint32_t* num_points);
int kzg_download_states(kzg_ctx* ctx, kz_body_state* out, uint32_t num_bodies);
/* counts[num_pairs], points[num_pairs * 4] */
int kzg_download_manifolds(kzg_ctx* ctx, uint32_t* counts, kz_contact_point* points);
/* hdr[4 * C] = body A slot, body B slot, num points, flags (1 sensor, 2 dynA, 4 dynB);
 * pts[(8 * c + k) * 44] = ContactPointConstraint fields, same layout as the
 * reference dump (oracle/ref_harness.cpp FillPoint). */
int kzg_download_constraints(kzg_ctx* ctx, uint32_t* hdr, float* pts);
/* RMS of the impulse increments of each velocity sweep of the last solve
 * (the PGS residual); returns the number of sweeps written. */
int kzg_get_residuals(kzg_ctx* ctx, double* out, uint32_t cap);
int kzg_get_timings(kzg_ctx* ctx, kzg_timings* out);

/* Standalone narrowphase over (shape, transform) pairs with the uploaded
 * shape table: xf = quaternion xyzw + translation, 7 floats.  Output:
 * counts[n], points[n * 4].  ms (optional) = device time of the kernels. */
int kzg_collide_batch(kzg_ctx* ctx, const uint32_t* shape_a, const uint32_t* shape_b,
                      const float* xf_a, const float* xf_b, uint32_t n, float max_separation,
                      uint32_t* counts, kz_contact_point* points, float* ms);

/* Host only (no GPU): the solve schedule kzg_solve would use for a contact
 * table, for tests and tuning.  info: 4 u32 per constraint (body A, body B,
 * first point, num points | sensor << 8), points numbered 0..P-1 in
 * constraint order.  Outputs per point: vel_level (velocity sweep level,
 * 0 = not swept) and ws_level (warm-start / bias level), levels counted from
 * 1 within the point's group; group[p] = its group.  stats = islands, groups,
 * deepest velocity group, deepest warm-start group.  Islands and the
 * first-fit batch order are those of PhysicsWorld::BuildIslands and
 * SolveIsland (src/World/PhysicsWorld.cpp:587-717). */
int kzg_schedule(uint32_t num_bodies, const uint8_t* dynamic, uint32_t num_constraints, const uint32_t* info,
                 int use_simd, uint32_t max_groups, uint32_t* vel_level, uint32_t* ws_level, uint32_t* group,
                 uint32_t* stats);

#ifdef __cplusplus
}
#endif

#endif /* KZG_H */
