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
    uint32_t allow_sleeping;     /* BodySettings::allowSleeping */
} kzg_body;

/* Built convex hulls: per shape the range of its vertices and triangles in
 * the pools (ConvexHullShape::Create output, src/Geometry/Shapes.cpp:331-471). */
typedef struct kzg_hull_range {
    uint32_t vert_offset, vert_count, tri_offset, tri_count;
} kzg_hull_range;

/* Per-step phase timings (milliseconds), host-device transfers excluded.
 * Device phases are measured with HIP events on the context's stream:
 * narrowphase (kernels, after the pair upload), the device part of the
 * build (constraint-table compaction and islands), integrate velocities,
 * solve (prepare, sweeps, write-back, cache store), the sweep kernel alone,
 * integrate positions and the constraint build kernel.  build_ms is host
 * compute of kzg_build_constraints once the constraint table has arrived:
 * the solve schedule (the build kernel runs meanwhile); the staging and
 * upload of the schedule are not included. */
typedef struct kzg_timings {
    float narrowphase_ms, build_ms, integrate_velocities_ms, solve_ms, integrate_positions_ms;
    float solve_kernel_ms;   /* the persistent sweep kernel alone */
    uint32_t velocity_levels, bias_levels, groups, islands;
    float build_kernel_ms;   /* the constraint build kernel alone */
    uint32_t cooperative_blocks;  /* workgroups that shared the largest group's levels */
    /* with cooperative_blocks > 1: that group's points in the warm-start/bias
     * and in the velocity sequences, and its sweep kernel alone */
    uint32_t cooperative_ws_points, cooperative_vel_points;
    float cooperative_kernel_ms;
    float schedule_ms;       /* the solve schedule alone (host, part of build_ms) */
    float build_device_ms;   /* device part of the build before the host schedule: compaction, islands */
    /* Inside schedule_ms: the warm-start / bias sequence (built on a helper
     * thread), and how long the velocity sequence then waited for it. */
    float schedule_ws_ms, schedule_wait_ms;
    /* Early warm start (SIMD solver, cache warm start, large worlds):
     * kzg_build_constraints returns once the warm-start schedule is ready and
     * kzg_solve sweeps the warm start while the host finishes the velocity
     * schedule.  early_warm_start_ms: the device time of that part of the
     * solve (inside solve_ms); schedule_velocity_after_ms: the host time the
     * velocity schedule took after the warm-start one was ready (inside
     * schedule_ms); overlap_ms: the device time (integrate velocities and the
     * early warm start) hidden under it, min of the two.  The step's device
     * and host phases then add up to narrowphase + build_device + build +
     * integrate_velocities + solve - overlap.  All zero otherwise. */
    float early_warm_start_ms, schedule_velocity_after_ms, overlap_ms;
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

/* Step 1's pair query on the device (DynamicAABBTree::QueryPairs +
 * BroadPhase::UpdatePairs, include/Kizuri/BroadPhase/BroadPhase.h:61-104,
 * src/BroadPhase/BroadPhase.cpp:368-381).  The caller keeps the tree as the
 * reference does (proxies created, moved and rebalanced on the host) and
 * passes its image: one record per node, leaves with child1 = 0xFFFFFFFF,
 * user = the proxy's body slot, height as the tree keeps it; the proxy
 * layers and masks indexed by node id.  The pairs come out in the
 * reference's order, the pairs its query loses included, as body slots, and
 * stay on the device for kzg_narrowphase_resident; *num_pairs receives
 * their count.  A tree whose ids, heights or user data break its invariants
 * is KZG_ERR_ARGUMENT. */
typedef struct kzg_bp_node {
    float fat_min[3];
    uint32_t child1;
    float fat_max[3];
    uint32_t child2;
    float tight_min[3];
    uint32_t user;
    float tight_max[3];
    int32_t height;
} kzg_bp_node;
int kzg_broadphase_pairs(kzg_ctx* ctx, const kzg_bp_node* nodes, uint32_t num_nodes, uint32_t root,
                         const uint32_t* layers, const uint32_t* masks, uint32_t num_layers, uint32_t* num_pairs);
/* Step 1's proxy moves on the device (UpdateBroadPhase, PhysicsWorld.cpp:333-345,
 * DynamicAABBTree::MoveProxy, src/BroadPhase/BroadPhase.cpp:58-75), against
 * the tree image the last kzg_broadphase_pairs / _update left on the device:
 * for every non-static body of the active list, its world bounds (the local
 * bounds registered with kzg_broadphase_set_proxies under its pose,
 * Shape::GetWorldBounds' arithmetic) become its leaf's tight box in that
 * image, and the bodies whose tight box left their leaf's fat box are
 * reported: *count of them, with out[7 * k] = the body's position in the
 * active list and out[7 * k + 1 .. 6] = its tight box (min, max), in no
 * particular order (at most cap are written).  The caller re-inserts those
 * leaves in active-list order, as MoveProxy would, and sends the changed
 * nodes with kzg_broadphase_pairs_update. */
int kzg_broadphase_set_proxies(kzg_ctx* ctx, const float* local_bounds, const uint32_t* proxies, uint32_t num_bodies);
int kzg_broadphase_move(kzg_ctx* ctx, float* out, uint32_t cap, uint32_t* count);
/* The pair query over the device's tree image after replacing n node
 * records (ids[k] gets records[k]); num_nodes must be the image's size and
 * root the tree's root.  Leaf tight boxes written by kzg_broadphase_move
 * stay unless a record replaces them. */
int kzg_broadphase_pairs_update(kzg_ctx* ctx, const uint32_t* ids, const kzg_bp_node* records, uint32_t n,
                                uint32_t num_nodes, uint32_t root, uint32_t* num_pairs);
/* The device's tree image (num_nodes records), e.g. to bring a host copy's
 * leaf tight boxes up to date after device moves. */
int kzg_broadphase_download(kzg_ctx* ctx, kzg_bp_node* out, uint32_t num_nodes);
/* Step 2 over the pairs the last kzg_broadphase_pairs left on the device. */
int kzg_narrowphase_resident(kzg_ctx* ctx);
/* The first n pairs the context holds (slot A, slot B each). */
int kzg_download_pairs(kzg_ctx* ctx, uint32_t* pairs, uint32_t n);

/* Step 3.  warm_start_mode: 0 = manifold cache (reference default),
 * 1 = cold start (the reference's KZ_NOWS switch). */
int kzg_build_constraints(kzg_ctx* ctx, int warm_start_mode);
int kzg_integrate_velocities(kzg_ctx* ctx, float dt);
/* Step 6. */
int kzg_solve(kzg_ctx* ctx, float dt);
int kzg_integrate_positions(kzg_ctx* ctx, float dt);

/* Step 8: PhysicsWorld::UpdateSleeping (PhysicsWorld.cpp:866-951), after
 * kzg_integrate_positions.  Sleep timers of the active bodies, union-find
 * over the last step's contact constraints with the island indices the
 * reference keeps on its bodies (active list positions; a sleeping body keeps
 * the index it had when it fell asleep), then every dynamic body of an
 * island whose smallest timer reached the threshold has its velocities
 * zeroed and leaves the active list (BodyManager::DeactivateBody,
 * src/Body/BodyManager.cpp:148-160: swap-remove), in the reference's loop
 * order.  A no-op in a world without sleeping (allow_sleeping = 0).
 *
 * With sleeping the context keeps the active list (BodyManager.h:63, 79):
 * kzg_narrowphase wakes sleeping bodies that a moving body touches
 * (BuildContactConstraints, PhysicsWorld.cpp:473-490: pushed onto the list in
 * constraint order), the integrations skip sleeping bodies (:760, :787,
 * :857), and the islands follow BuildIslands (:587-664) over the active list:
 * compact indices for its dynamic bodies, while a sleeping body's contacts
 * are routed through the stale index it keeps, as in the reference. */
int kzg_update_sleeping(kzg_ctx* ctx, float dt);
/* The active body list in order, as body slots: *count receives its
 * length; at most cap entries are written. */
int kzg_get_active_bodies(kzg_ctx* ctx, uint32_t* out, uint32_t cap, uint32_t* count);
/* Replace the active list: the listed bodies are active in this order, every
 * other body sleeps with zero velocities; sleep timers are zero and island
 * indices unset.  That is the state that creating the bodies in slot order
 * and deactivating the unlisted ones leaves (BodyManager::CreateBody /
 * DeactivateBody, BodyManager.cpp:53-59, 148-160; the snapshot loader,
 * src/Serialization/Serializer.cpp:230-234), when the list is in slot order.
 * kzg_upload_bodies resets the list to every body in slot order. */
int kzg_set_active_bodies(kzg_ctx* ctx, const uint32_t* list, uint32_t count);
/* WorldSettings::sleepTimeThreshold / sleepLinearThreshold /
 * sleepAngularThreshold (include/Kizuri/World/PhysicsWorld.h:25-27);
 * defaults 0.5 s, 0.05 m/s, 0.1 rad/s. */
int kzg_set_sleep_thresholds(kzg_ctx* ctx, float time, float linear, float angular);

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
/* Natural-map residual of the iterate the last kzg_solve left (call before
 * kzg_integrate_positions), over the non-sensor contact points, with the
 * bodies' velocities and the anchors the velocity sweep uses (rA = qA
 * localAnchorA).  Per point: normal complementarity min(lambda_n, m_n
 * (vRel . n - bias)); friction lambda_t - clamp(lambda_t - m_t (vRel . t),
 * +-mu lambda_n) for t1 and t2; cone max(0, |lambda_t| - mu lambda_n) (a part
 * of the friction term).  bias = restitution bias (+ position bias with
 * velocity bias).  out[4] = RMS of normal + friction, of normal, of
 * friction, of cone, in N s.  No reference counterpart (the reference
 * computes no residual). */
int kzg_get_natural_residual(kzg_ctx* ctx, double* out);
int kzg_get_timings(kzg_ctx* ctx, kzg_timings* out);

/* The warm-start cache (the contact cache of PhysicsWorld.cpp:405-452),
 * for snapshots: capacity in slots (a power of two) and occupied slots.
 * Its image is KZG_CACHE_SLOT_BYTES bytes per slot in four arrays: keys
 * (u64 MakePairKey, 0 = empty), point counts (u32), points (4 float4 per
 * slot: pointOnA and normal impulse), tangent impulses (4 float2 per slot).
 * kzg_upload_cache replaces the cache with such an image. */
#define KZG_CACHE_SLOT_BYTES 108
int kzg_cache_info(kzg_ctx* ctx, uint32_t* capacity, uint32_t* used);
int kzg_download_cache(kzg_ctx* ctx, void* out, uint64_t bytes);
int kzg_upload_cache(kzg_ctx* ctx, const void* in, uint64_t bytes, uint32_t capacity, uint32_t used);

/* Standalone narrowphase over (shape, transform) pairs with the uploaded
 * shape table: xf = quaternion xyzw + translation, 7 floats.  Output:
 * counts[n], points[n * 4].  ms (optional) = device time of the kernels. */
int kzg_collide_batch(kzg_ctx* ctx, const uint32_t* shape_a, const uint32_t* shape_b,
                      const float* xf_a, const float* xf_b, uint32_t n, float max_separation,
                      uint32_t* counts, kz_contact_point* points, float* ms);

/* kzg_collide_batch with every pointer in device memory and no device
 * allocation: no counterpart in the reference.  It is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  shape_a, shape_b: n u32 shape indices
 * into the uploaded table; xf_a, xf_b: n x 7 floats (quaternion xyzw +
 * translation); outputs counts[n] and points[n * 4] as kzg_collide_batch
 * writes them.  The narrowphase's own buffers (work lists, GJK records,
 * manifold slots) come from a scratch array the library holds on each GPU
 * for the life of the process (about 9 MB), which bounds n to
 * KZG_NO_ALLOC_MAX_PAIRS; a larger n is KZG_ERR_ARGUMENT.  Every call on
 * one GPU shares that GPU's array, from any context: calls that overlap on
 * one GPU race.  The call is synchronous on the context's stream.  A shape
 * index outside the table gives that pair no contact and the call returns
 * KZG_ERR_ARGUMENT. */
#define KZG_NO_ALLOC_MAX_PAIRS 16384
int kzg_collide_batch_device(kzg_ctx* ctx, const uint32_t* shape_a, const uint32_t* shape_b,
                             const float* xf_a, const float* xf_b, uint32_t n, float max_separation,
                             uint32_t* counts, kz_contact_point* points);

/* Warm-start modes of kzg_build_constraints / kzg_set_warm_start.  The
 * reference initialises accumulated impulses from its manifold cache
 * (PhysicsWorld.cpp:492-530) or, under KZ_NOWS, with zeros (:493).  The
 * learned modes add the contact-point network's correction (kzg_mlp_*) to
 * the cache's impulses: at every point, or only where the cache has no
 * match.  The sums are clamped as Prepare clamps cached impulses
 * (ContactSolver.cpp:107-110).  KZG_WARM_NO_STORE, or-ed into a mode,
 * makes the following kzg_solve leave the cache untouched (for comparing
 * modes on one state); KZG_WARM_NO_COLLECT makes it append no training rows
 * while collection is on (kzg_mlp_collect). */
enum kzg_warm_start {
    KZG_WARM_CACHE = 0,
    KZG_WARM_COLD = 1,
    KZG_WARM_LEARNED = 2,
    KZG_WARM_LEARNED_MISS = 3,
    KZG_WARM_NO_STORE = 0x100,
    KZG_WARM_NO_COLLECT = 0x200
};
/* Mode kzg_step uses (default KZG_WARM_CACHE). */
int kzg_set_warm_start(kzg_ctx* ctx, int mode);
/* Override SolverSettings::velocityIterations / positionIterations. */
int kzg_set_iterations(kzg_ctx* ctx, uint32_t velocity, uint32_t position);
/* Workgroups that share the levels of the largest group of islands (the
 * one holding the largest island) in the sweeps, with a device-wide barrier
 * between levels: 0 = chosen from the level widths (default), 1 = one
 * workgroup per group, n = n workgroups (at most 64).  Results are the
 * same whatever the setting: a level's points share no dynamic body. */
int kzg_set_cooperative(kzg_ctx* ctx, uint32_t blocks);
/* Where the workgroups that share the largest group's levels run.
 * KZG_COOP_ONE_XCD (default): all on one XCD, which gives the levels a
 * barrier through that XCD's L2 with no cache write-back (the launch holds
 * 8 workgroups per participant and one XCD's take part; at most 32).
 * KZG_COOP_SPREAD: anywhere on the device, with agent-scope fences between
 * levels.  Results are the same either way. */
#define KZG_COOP_ONE_XCD 0u
#define KZG_COOP_SPREAD 1u
int kzg_set_cooperative_placement(kzg_ctx* ctx, uint32_t placement);
/* Where the velocity schedule of an early warm start (kzg_set_early_warm_start)
 * is bucketed by level: 0 = on the device when one group of islands holds at
 * least half the points (default), 1 = on the host.  The device takes the
 * host's level walk per point and builds the level offsets, unit bounds and
 * record positions in parallel; outside the largest group the units of one
 * level then sit in another order, which changes no result (they share no
 * dynamic body). */
int kzg_set_device_bucketing(kzg_ctx* ctx, uint32_t mode);
/* Bound on the polls of one wait between levels of a cooperative sweep
 * (0 = default, about a second).  A sweep whose wait runs out stops early
 * on every workgroup, and the next call that waits for the device (build,
 * state download, timings) returns KZG_ERR_DEVICE. */
int kzg_set_spin_limit(kzg_ctx* ctx, uint32_t polls);
/* Most points of one body pair a sweep thread applies in a row (a unit), in
 * the velocity and in the warm-start / bias sweeps: 0 = default (2 each),
 * 1..4.  A point whose dynamic bodies were last updated by the same point of
 * the same pair joins that point's unit instead of starting a new level,
 * which shortens the level chain; each unit costs its points' arithmetic in
 * series.  Results are the same whatever the setting. */
int kzg_set_sweep_units(kzg_ctx* ctx, uint32_t velocity, uint32_t warm_start);
/* Waves of the narrowphase's EPA / clipping kernel (one pair per wave at a
 * time, the next taken from a shared counter): 0 = auto (16 per compute
 * unit, at most one per pair).  Results are the same for every count. */
int kzg_set_narrowphase_waves(kzg_ctx* ctx, uint32_t waves);

/* Islands (BuildIslands, PhysicsWorld.cpp:587-664) are a union-find on the
 * device: in one workgroup's LDS for worlds of at most `max_bodies` bodies
 * (0 = default, 36,864, also the ceiling), else in global memory with
 * atomics.  1 sends every world to the global-memory version.  Results are
 * the same either way (roots are each component's smallest slot). */
int kzg_set_island_lds_limit(kzg_ctx* ctx, uint32_t max_bodies);

/* Early warm start (no reference counterpart: an overlap of host and device
 * work inside PhysicsWorld::StepFixed's order).  With the SIMD solver and the
 * cache warm start, kzg_build_constraints may return once the warm-start
 * schedule is built, and kzg_solve then sweeps the warm start and the bias
 * (position) iterations on the device while the host finishes the velocity
 * (first-fit) schedule.  mode: 0 = auto (worlds of 16,384 or more contact
 * points), 1 = off, 2 = always.  Results are the same either way. */
int kzg_set_early_warm_start(kzg_ctx* ctx, uint32_t mode);

/* ---- Learned warm start (no reference counterpart) ---------------------
 * A per-contact-point MLP 16 -> 64 -> 64 -> 3 with ReLU hidden layers,
 * BF16 weights and FP32 accumulation on the matrix cores; it replaces the
 * impulse initialisation at PhysicsWorld.cpp:514-530 and is trained on the
 * impulses the solver stores at :568-581.
 * Feature row, 16 floats per point, all network inputs: 0 normal relative
 * velocity, 1-2 relative velocity along t1 / t2, 3-5 normal, 6 depth, 7-8
 * inverse masses A / B, 9-10 (r x n).I^-1(r x n) for A / B, 11-12 the mean
 * of that along t1 and t2 for A / B, 13 cache hit (0/1), 14-15 the loads of
 * bodies A / B with KZG_FEATURES_LOAD (kzg_set_learned_features), else 0.
 * Inputs are normalised with a per-feature mean and inverse standard
 * deviation (a constant feature gets inverse deviation 0).  Outputs y are
 * impulse corrections in units of the point's impulse scale s = normal mass
 * * |gravity| * dt on top of the cache's clamped impulses c (zero where the
 * cache has no match): lambda_n = max(c_n + y0 s, 0),
 * lambda_t = clamp(c_t + y1,2 s, +-mu lambda_n).  KZG_WARM_LEARNED applies
 * the correction everywhere, KZG_WARM_LEARNED_MISS only where c is zero by
 * a miss.
 * Parameters (FP32 master copy), in this order: W1[64][16], b1[64],
 * W2[64][64], b2[64], W3[3][64], b3[3] (row = output unit). */
#define KZG_MLP_PARAMS 5443
#define KZG_MLP_FEATURES 16
#define KZG_MLP_FEATURE_ROW 16
#define KZG_MLP_SAMPLE_ROW 20   /* feature row + target correction y (3) + 0 */
/* He-uniform weights from a SplitMix64 seed, zero biases, Adam reset. */
int kzg_mlp_init(kzg_ctx* ctx, uint64_t seed);
/* Which optional features kzg_solve writes into features 14-15 (0: none,
 * the network sees the 14 contact-frame features alone).
 * KZG_FEATURES_LOAD: the load of each of the point's bodies, the sum over
 * all its contact points of the cached normal impulse clamped at 0, in units
 * of the body's weight impulse m |gravity| dt (0 for a static or kinematic
 * body); summed in fixed point, so the same in any order.  It tells the
 * network how much of the stack a body carries, which no contact-frame
 * feature shows. */
#define KZG_FEATURES_LOAD 1u
int kzg_set_learned_features(kzg_ctx* ctx, uint32_t flags);
int kzg_mlp_set_params(kzg_ctx* ctx, const float* params, uint32_t n);
int kzg_mlp_get_params(kzg_ctx* ctx, float* params, uint32_t n);
int kzg_mlp_set_normalization(kzg_ctx* ctx, const float* mean, const float* inv_std);
int kzg_mlp_get_normalization(kzg_ctx* ctx, float* mean, float* inv_std);
/* Training data: every kzg_solve appends, per point, its feature row and
 * (final accumulated impulses - c) / s to a device ring of `capacity` rows
 * (0 switches collection off; the call empties the ring). */
int kzg_mlp_collect(kzg_ctx* ctx, uint32_t capacity);
int kzg_mlp_num_samples(kzg_ctx* ctx, uint64_t* written, uint32_t* held);
/* Rows first .. first+count-1 of the ring, KZG_MLP_SAMPLE_ROW floats each. */
int kzg_mlp_get_samples(kzg_ctx* ctx, uint32_t first, uint32_t count, float* out);
/* Feature mean / inverse standard deviation from up to 65536 ring rows. */
int kzg_mlp_fit_normalization(kzg_ctx* ctx);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) on minibatches of `batch`
 * rows drawn from the ring by a hash of (seed, step, row); loss = mean
 * squared error over the batch and the 3 outputs; losses[steps] optional. */
int kzg_mlp_train(kzg_ctx* ctx, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float* losses);
/* Training targets clamped to [-clip, clip] (units of s) by kzg_mlp_train;
 * 0 (the default) trains on them as stored.  A few separated pairs whose
 * clipped manifolds carry impulses of hundreds of s otherwise dominate the
 * squared error.  No reference counterpart (the learned warm start is new). */
int kzg_mlp_set_target_clip(kzg_ctx* ctx, float clip);
/* Inference on feature rows (n x 16 floats): out = (y0, y1, y2, 0) per row. */
int kzg_mlp_forward(kzg_ctx* ctx, const float* features, uint32_t n, float* out);
/* Gradient of the mean squared error over sample rows (n x 20 floats) with
 * respect to the parameters (KZG_MLP_PARAMS floats); loss = summed squared error. */
int kzg_mlp_gradients(kzg_ctx* ctx, const float* samples, uint32_t n, float* grad, float* loss);

/* Host only (no GPU): the solve schedule kzg_solve would use for a contact
 * table, for tests and tuning.  info: 4 u32 per constraint (body A, body B,
 * first point, num points | sensor << 8), points numbered 0..P-1 in
 * constraint order.  flags: KZG_SCHEDULE_SIMD for the first-fit velocity
 * order (else the constraint order), KZG_SCHEDULE_UNITS to join points into
 * multi-point units with kzg_solve's default caps (else one point per
 * unit), KZG_SCHEDULE_DEFER to build it as kzg_build_constraints does for
 * an early warm start (warm-start order first, velocity order finished on
 * the helper thread; the result is the same).  Outputs
 * per point (each optional): vel_level (velocity sweep level, 0 = not
 * swept) and ws_level (warm-start / bias level), levels counted from 1
 * within the point's group; group[p] = its group; vel_unit / ws_unit = the
 * unit holding the point (units numbered in sweep order; the points of a
 * unit run in point order).  stats (6) = islands, groups, deepest velocity
 * group, deepest warm-start group, velocity units, warm-start units.
 * Islands and the first-fit batch order are those of
 * PhysicsWorld::BuildIslands and SolveIsland
 * (src/World/PhysicsWorld.cpp:587-717). */
#define KZG_SCHEDULE_SIMD 1u
#define KZG_SCHEDULE_UNITS 2u
#define KZG_SCHEDULE_DEFER 4u
int kzg_schedule(uint32_t num_bodies, const uint8_t* dynamic, uint32_t num_constraints, const uint32_t* info,
                 uint32_t flags, uint32_t max_groups, uint32_t* vel_level, uint32_t* ws_level, uint32_t* group,
                 uint32_t* vel_unit, uint32_t* ws_unit, uint32_t* stats);

#ifdef __cplusplus
}
#endif

#endif /* KZG_H */
